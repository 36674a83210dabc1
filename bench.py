#!/usr/bin/env python3
"""Benchmark: RBRP column ID time-to-tolerance on one MI355X.

Workload (one "step" = one complete ID to tolerance, X included): A = U diag(s) V^T
with Haar U (50000 x 4000), Haar V (4000 x 4000) and s_i = exp(-i/100), FP64,
block size b = 64, tol = 1e-8 (relative Frobenius, non-squared).  A is generated
on the device and is resident in HBM before the timed region starts.

Prints one JSON line: value = IDs per second (higher is better), the roofline
of the dominant kernel (per-launch hipEvent durations measured by the library
over the timed region, on the stream the kernels run on) and the CPU baseline
(the NumPy/SciPy oracle on the same matrix and Philox stream, timed on this
host).  Further legs in the same line:
  mixed    -- FP32-stored residual vs FP64 at the same tolerance (1e-6) on the
              50000 x 4000 exp-decay and the 100000 x 2000 Gaussian-mixture matrices;
  batched  -- 4096 Gaussian-kernel blocks (m ~ U[64,1024], n ~ U[64,512]), b = 16,
              tol = 1e-7: one persistent launch vs one launch sequence per problem;
  scale    -- 1,000,000 x 1024 FP32-stored matrix (s_i = 1/i, 500 Cauchy rows),
              b = 128, tol = 1e-5: update GB/s and projection MFMA utilisation;
  comparators -- RandCPQR_OS / RandLUPP_OS (and RandCPQR / RandLUPP, without
              oversampling) at the rank RBRP found;
  cpqr     -- the full-matrix CPQR ID at ranks of the RBRP trace (time model,
              errors, and the Householder steps' HBM GB/s).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "RBRP ID time-to-tol, mixed vs FP64 on 1 GPU at equal rel. error; batched IDs/sec"
PEAK_HBM_GBS = 8000.0
PEAK_FP64_TFLOPS = 78.6
PEAK_FP32_TFLOPS = 157.3


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--m", type=int, default=50000)
    ap.add_argument("--n", type=int, default=4000)
    ap.add_argument("--rate", type=float, default=100.0, help="s_i = exp(-i/rate)")
    ap.add_argument("--b", type=int, default=64)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--matrix-seed", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-comparators", action="store_true",
                    help="skip the RandCPQR_OS / RandLUPP_OS comparison runs")
    ap.add_argument("--profile-steps", action="store_true",
                    help="print a line per step (for profiler runs)")
    ap.add_argument("--traffic-probe", action="store_true",
                    help="run a 1 GiB device copy and one workload (--probe-what), print nothing "
                         "else (for rocprofv3 --pmc FETCH_SIZE / WRITE_SIZE passes); needs "
                         "--load-matrix, so no solver library runs in the counted process")
    ap.add_argument("--probe-what", default="colid", choices=["colid", "cpqr"],
                    help="workload of --traffic-probe: the FP64 RBRP ID or the full-matrix CPQR ID")
    ap.add_argument("--save-matrix", default="",
                    help="write the generated matrix (column-major .npy) and exit")
    ap.add_argument("--load-matrix", default="",
                    help="read the matrix from this .npy instead of generating it")
    ap.add_argument("--cpu-reps", type=int, default=3,
                    help="oracle IDs of the benchmark matrix timed for the CPU baseline (median)")
    ap.add_argument("--legs", default="main,comparators,cpqr,mixed,batched,scale,ceilings,sampler,cpu",
                    help="comma list of legs to run (main = the timed FP64 IDs)")
    ap.add_argument("--no-mixed", action="store_true", help="skip the mixed-precision legs")
    ap.add_argument("--all-phase-timing", action="store_true",
                    help="record events around every phase inside the timed IDs (default: "
                         "projection and update only; the breakdown comes from one extra ID)")
    ap.add_argument("--no-batched", action="store_true", help="skip the batched leg")
    ap.add_argument("--scale-m", type=int, default=1000000,
                    help="rows of the 1024-column FP32 matrix of the scale leg")
    ap.add_argument("--batch-count", type=int, default=4096)
    ap.add_argument("--no-batch-loop", action="store_true",
                    help="skip the per-problem launch loop of the batched leg")
    ap.add_argument("--batch-cpu-sample", type=int, default=48,
                    help="problems of the batch timed with the CPU oracle")
    args = ap.parse_args()
    if args.traffic_probe and not args.load_matrix:
        # the device generator factors its Haar matrices with torch.linalg.qr
        # (rocSOLVER), which has crashed inside counter-collection processes:
        # the matrix must come from a file written by a separate process
        ap.error("--traffic-probe needs --load-matrix (write it first with --save-matrix in a "
                 "process of its own)")
    legs = set(x.strip() for x in args.legs.split(",") if x.strip())
    args.no_comparators = args.no_comparators or "comparators" not in legs
    args.no_cpqr = "cpqr" not in legs
    args.no_mixed = args.no_mixed or "mixed" not in legs
    args.no_batched = args.no_batched or "batched" not in legs
    args.no_cpu = args.no_cpu or "cpu" not in legs
    args.no_scale = "scale" not in legs
    args.no_ceilings = "ceilings" not in legs
    args.no_sampler = "sampler" not in legs
    args.no_main = "main" not in legs
    return args


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "traffic_k_gemm_upd64_fp64_50000x4000.json")


def measured_traffic(args):
    """HBM bytes per launch of the update kernel from the committed rocprofv3 --pmc
    summary of this workload (FETCH_SIZE x its copy-kernel factor + WRITE_SIZE), or None."""
    try:
        with open(TRAFFIC_FILE) as f:
            t = json.load(f)
    except (OSError, ValueError):
        return None
    if (t.get("m"), t.get("n"), t.get("b"), t.get("tol")) != (args.m, args.n, args.b, args.tol):
        return None
    return t.get("bytes_per_launch")


def _sync():
    import torch
    torch.cuda.synchronize()


def id_rel_error(A, res):
    """||A(:, rd) - A(:, J) X||_F / ||A||_F of a returned ID, on the device."""
    import torch
    n = A.shape[1]
    if res.k >= n:
        return 0.0
    perm = res.perm.long()
    E = A[:, perm[res.k:]] - A[:, perm[: res.k]] @ res.X
    e = (torch.linalg.norm(E) / torch.linalg.norm(A)).item()
    del E
    return e


def time_colid(A, b, tol, seed, precision, h, reps):
    """One warm-up and `reps` timed IDs; returns (last result, seconds per ID)."""
    from robust_blockwise_random_pivoting_amd import rbrp_colid
    rbrp_colid(A, b=b, tol=tol, seed=seed, precision=precision, handle=h)
    _sync()
    t0 = time.perf_counter()
    for _ in range(reps):
        res = rbrp_colid(A, b=b, tol=tol, seed=seed, precision=precision, handle=h)
    _sync()
    return res, (time.perf_counter() - t0) / reps


def mixed_leg(A2, args):
    """Mixed precision (FP32-stored residual, FP32 MFMA) vs FP64 at the same tolerance:
    the 50000 x 4000 exp-decay matrix at tol 1e-6 (the tightest tolerance FP32
    storage resolves on it) and the 100000 x 2000 Gaussian-mixture matrix at 1e-6."""
    import torch
    from robust_blockwise_random_pivoting_amd import Handle, rbrp_colid
    from robust_blockwise_random_pivoting_amd.matrices import gmm_rows_torch
    out = {}
    with Handle() as h:
        cases = [("exp_50000x4000_b64_tol1e-6", lambda: A2, 64, 1e-6),
                 ("gmm_100000x2000_b64_tol1e-6",
                  lambda: gmm_rows_torch(100000, 2000, 7, components=20), 64, 1e-6)]
        for label, gen, b, tol in cases:
            A = gen()
            _sync()
            r64, t64 = time_colid(A, b, tol, args.seed, "fp64", h, 2)
            r32, t32 = time_colid(A, b, tol, args.seed, "fp32", h, 2)
            e64, e32 = id_rel_error(A, r64), id_rel_error(A, r32)
            out[label] = {"fp64_ms": t64 * 1e3, "fp32_ms": t32 * 1e3, "speedup": t64 / t32,
                          "k_fp64": r64.k, "k_fp32": r32.k, "rel_error_fp64": e64,
                          "rel_error_fp32": e32,
                          "rel_error_diff": abs(e32 - e64) / e64 if e64 > 0 else 0.0,
                          "reported_error_fp32": r32.err_final}
            # where the time goes in each form (one more ID each, every phase timed)
            for prec in ("fp64", "fp32"):
                h.set_timing(1)
                h.reset_stats()
                rbrp_colid(A, b=b, tol=tol, seed=args.seed, precision=prec, handle=h)
                _sync()
                out[label][f"phase_ms_{prec}"] = {k: v["ms"] for k, v in h.stats().items()}
                h.set_timing(0)
            print(f"[bench] mixed {label}: fp64 {t64 * 1e3:.1f} ms k={r64.k} err={e64:.3e} | "
                  f"fp32 {t32 * 1e3:.1f} ms k={r32.k} err={e32:.3e}", file=sys.stderr, flush=True)
            if A is not A2:
                del A
                torch.cuda.empty_cache()
    return out


def scale_leg(args):
    """Mixed form at scale: 1,000,000 x 1024 FP32-stored matrix, s_i = 1/i (Haar U, V),
    500 rows replaced by standard-Cauchy rows; b = 128, tol = 1e-5.  The whole
    matrix (4.1 GB) is resident in HBM.  Reports the wall time of the ID, the
    per-phase device time, the update's HBM GB/s and the projection's FP32 MFMA
    utilisation, and the FP64 form on the widened matrix for comparison."""
    import torch
    from robust_blockwise_random_pivoting_amd import Handle, rbrp_colid
    from robust_blockwise_random_pivoting_amd.matrices import poly_outlier_torch
    m, n, b, tol = args.scale_m, 1024, 128, 1e-5
    A = poly_outlier_torch(m, n, seed=5)
    _sync()
    print(f"[bench] scale: {m}x{n} FP32 matrix ready", file=sys.stderr, flush=True)
    out = {"m": m, "n": n, "b": b, "tol": tol, "stored": "fp32"}
    with Handle() as h:
        r32, t32 = time_colid(A, b, tol, args.seed, "fp32", h, 2)
        h.set_timing(True)
        h.reset_stats()
        rbrp_colid(A, b=b, tol=tol, seed=args.seed, precision="fp32", handle=h)
        _sync()
        st = h.stats()
        h.set_timing(False)
        upd, prj = st["update"], st["project"]
        out.update({
            "fp32_ms": t32 * 1e3, "k": r32.k, "blocks": len(r32.ranks),
            "tracked_error": r32.err_final,
            "phase_ms": {k: v["ms"] for k, v in st.items()},
            "update_GBs": upd["bytes"] / (upd["ms"] * 1e-3) / 1e9 if upd["ms"] > 0 else None,
            "update_hbm_frac": (upd["bytes"] / (upd["ms"] * 1e-3) / 1e9 / PEAK_HBM_GBS
                                if upd["ms"] > 0 else None),
            "project_TFLOPs": prj["flops"] / (prj["ms"] * 1e-3) / 1e12 if prj["ms"] > 0 else None,
            "project_mfma_frac_fp32": (prj["flops"] / (prj["ms"] * 1e-3) / 1e12 / PEAK_FP32_TFLOPS
                                       if prj["ms"] > 0 else None),
            "project_GBs": prj["bytes"] / (prj["ms"] * 1e-3) / 1e9 if prj["ms"] > 0 else None,
        })
        print(f"[bench] scale fp32: {t32 * 1e3:.1f} ms k={r32.k} update "
              f"{out['update_GBs']:.0f} GB/s project {out['project_GBs']:.0f} GB/s",
              file=sys.stderr, flush=True)
        A64 = A.double()
        r64, t64 = time_colid(A64, b, tol, args.seed, "fp64", h, 1)
        out.update({"fp64_ms": t64 * 1e3, "k_fp64": r64.k, "speedup": t64 / t32,
                    "tracked_error_fp64": r64.err_final,
                    "rel_error_fp32": id_rel_error(A64, r32),
                    "rel_error_fp64": id_rel_error(A64, r64)})
        print(f"[bench] scale fp64: {t64 * 1e3:.1f} ms k={r64.k}", file=sys.stderr, flush=True)
        del A64
    del A
    torch.cuda.empty_cache()
    return out


def batched_leg(args):
    """4096 Gaussian-kernel blocks (m ~ U[64,1024], n ~ U[64,512]), b = 16, tol = 1e-7:
    one persistent launch (autotuned, replayed from its hipGraph) vs one launch
    sequence per problem, and the CPU oracle on a sample of the same problems."""
    import torch
    from robust_blockwise_random_pivoting_amd import Batch, Handle, batched_launch_shape, batched_tune
    from robust_blockwise_random_pivoting_amd.matrices import kernel_blocks_torch
    nprob = args.batch_count
    As = kernel_blocks_torch(nprob, seed=4, m_range=(64, 1024), n_range=(64, 512))
    _sync()
    print(f"[bench] batched: {nprob} kernel blocks ready", file=sys.stderr, flush=True)
    b, tol = 16, 1e-7
    seeds = list(range(nprob))
    with Handle() as h:
        wg = batched_tune(As, b=b, tol=tol, seeds=seeds, handle=h)
        _, tile = batched_launch_shape(As, b=b, tol=tol, handle=h)
        pb = Batch(As, b=b, tol=tol, seeds=seeds)
        st = pb.run("persistent", h)                       # warm-up: records the graph
        reps = 3
        t0 = time.perf_counter()
        for _ in range(reps):
            st = pb.run("persistent", h)
        tp = (time.perf_counter() - t0) / reps
        print(f"[bench] batched persistent: {tp * 1e3:.1f} ms ({wg} WG/CU, {tile} row tiles)",
              file=sys.stderr, flush=True)
        if args.no_batch_loop:
            lb, stl, tl = None, np.zeros(nprob, np.int32), float("nan")
        else:
            lb = Batch(As, b=b, tol=tol, seeds=seeds)
            t0 = time.perf_counter()
            stl = lb.run("loop", h)
            tl = time.perf_counter() - t0
            print(f"[bench] batched loop: {tl * 1e3:.1f} ms", file=sys.stderr, flush=True)
    rp = pb.results()
    same = None if lb is None else all(
        a.k == c.k and torch.equal(a.perm, c.perm) for a, c in zip(rp, lb.results()))
    ks = np.array([r.k for r in rp])
    errs = np.array([r.err_final for r in rp])
    out = {"problems": nprob, "b": b, "tol": tol, "persistent_ms": tp * 1e3,
           "persistent_ids_per_s": nprob / tp, "loop_ms": tl * 1e3, "loop_ids_per_s": nprob / tl,
           "speedup_vs_loop": tl / tp, "wg_per_cu": wg, "update_row_tiles": tile,
           "status_nonzero": int((st != 0).sum()),
           "loop_status_nonzero": int((stl != 0).sum()), "same_as_loop": same,
           "k_mean": float(ks.mean()), "k_max": int(ks.max()),
           "max_tracked_error": float(errs.max())}
    if not args.no_cpu and args.batch_cpu_sample > 0:
        from oracle import rbrp_colid as oracle_colid
        q = min(args.batch_cpu_sample, nprob)
        hostA = [As[i].cpu().numpy() for i in range(q)]
        t0 = time.perf_counter()
        agree = 0
        for i in range(q):
            r = oracle_colid(hostA[i], b, tol, seed=seeds[i])
            agree += int(r.k == rp[i].k and np.array_equal(r.perm[: r.k],
                                                          rp[i].perm[: r.k].cpu().numpy()))
        dt = time.perf_counter() - t0
        out["cpu_baseline"] = {"value": q / dt, "unit": "IDs/s", "kind": "port",
                               "sample": f"first {q} problems of the batch, one after another",
                               "skeletons_equal_to_gpu": f"{agree}/{q}"}
    del As
    torch.cuda.empty_cache()
    return out


def ceilings_leg():
    """Measured ceilings of this part (committed micro-kernels of the test-hook
    library): a streaming read and copy (nontemporal 16-byte accesses), an
    in-place read-modify-write sweep in the trailing update's tile order (the
    ceiling of that access pattern), and loops of independent FP64 / FP32
    16x16x4 matrix-core instructions."""
    import ctypes
    import torch
    from robust_blockwise_random_pivoting_amd import _capi
    L = _capi.testing_lib()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    out = {}
    for kind, name, nbytes, unit, peak in ((0, "read", 2 << 30, "GB/s", PEAK_HBM_GBS),
                                           (1, "copy", 1 << 30, "GB/s", PEAK_HBM_GBS),
                                           (4, "update_sweep", 1200 << 20, "GB/s", PEAK_HBM_GBS),
                                           (2, "mfma_f64", 1 << 20, "TFLOP/s", PEAK_FP64_TFLOPS),
                                           (3, "mfma_f32", 1 << 20, "TFLOP/s", PEAK_FP32_TFLOPS)):
        ms, rate = ctypes.c_double(), ctypes.c_double()
        _capi.check("rbrp_test_ceiling", L.rbrp_test_ceiling(st, kind, nbytes, 10, ctypes.byref(ms),
                                                             ctypes.byref(rate)))
        out[name] = {"rate": rate.value, "unit": unit, "peak": peak, "frac": rate.value / peak,
                     "ms": ms.value}
        print(f"[bench] ceiling {name}: {rate.value:.1f} {unit}", file=sys.stderr, flush=True)
    return out


def sampler_leg(args):
    """The weighted sampler at a candidate range of 2^26 (the multi-workgroup form):
    one block of b = 64 draws from 2^26 random weights, device time per call from
    events over 10 calls.  Bytes moved: the weights once (8 per candidate) plus
    12 per 4096-candidate tile; the SURVEY's model charges 16 per candidate
    (weights read and a CDF written: this form never stores the CDF)."""
    import ctypes
    import torch
    from robust_blockwise_random_pivoting_amd import _capi
    L = _capi.testing_lib()
    n, b = 1 << 26, 64
    gen = torch.Generator(device="cuda")
    gen.manual_seed(3)
    w = torch.rand(n, dtype=torch.float64, device="cuda", generator=gen)
    _sync()
    ms = ctypes.c_double()
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    _capi.check("rbrp_test_sample_time", L.rbrp_test_sample_time(
        st, ctypes.c_void_p(w.data_ptr()), n, b, args.seed, 10, ctypes.byref(ms)))
    t = ms.value * 1e-3
    moved = 8.0 * n + 12.0 * (n // 4096)
    out = {"n": n, "b": b, "ms": ms.value, "bytes_moved": moved, "GBs": moved / t / 1e9,
           "hbm_frac": moved / t / 1e9 / PEAK_HBM_GBS, "model_bytes": 16.0 * n,
           "model_GBs": 16.0 * n / t / 1e9, "model_frac": 16.0 * n / t / 1e9 / PEAK_HBM_GBS}
    print(f"[bench] sampler n=2^26: {ms.value:.3f} ms, {out['GBs']:.0f} GB/s moved",
          file=sys.stderr, flush=True)
    del w
    torch.cuda.empty_cache()
    return out


def _blas_threads(default):
    try:
        from threadpoolctl import threadpool_info
        info = threadpool_info()
        return max([t.get("num_threads", 1) for t in info if t.get("user_api") == "blas"] or [default])
    except Exception:
        return default


def cpu_baseline(A_host, args, k_gpu, perm_gpu):
    """The oracle (NumPy/SciPy LAPACK, multi-threaded BLAS) on the same matrix and
    stream: the median of args.cpu_reps IDs of the benchmark matrix (the reported
    rate); plus config 1 (1000 x 1000, s_i = exp(-i/20), b = 40, tol 1e-6) as the
    median of 10 with every BLAS thread and with one thread."""
    import statistics
    from oracle import id_error, rbrp_colid as oracle_colid
    from robust_blockwise_random_pivoting_amd.matrices import exp_spectrum, spectrum_matrix
    try:
        affinity = sorted(os.sched_getaffinity(0))
    except Exception:
        affinity = list(range(os.cpu_count() or 1))
    cores = len(affinity)
    threads = _blas_threads(int(os.environ.get("OMP_NUM_THREADS", cores)))
    times, r = [], None
    for _ in range(max(1, args.cpu_reps)):
        t0 = time.perf_counter()
        r = oracle_colid(A_host, args.b, args.tol, seed=args.seed)
        times.append(time.perf_counter() - t0)
    dt = statistics.median(times)
    same = bool(r.k == k_gpu and np.array_equal(r.perm[: r.k], perm_gpu[:k_gpu]))
    # config 1, all threads and one thread
    A1 = spectrum_matrix(1000, 1000, exp_spectrum(1000, 20.0), 1)
    c1 = {}
    try:
        from threadpoolctl import threadpool_limits
    except Exception:
        threadpool_limits = None
    for label, lim in (("all_threads", None), ("one_thread", 1)):
        ts = []
        for _ in range(10):
            if threadpool_limits is not None and lim is not None:
                with threadpool_limits(limits=lim):
                    t0 = time.perf_counter()
                    r1 = oracle_colid(A1, 40, 1e-6, seed=42)
                    ts.append(time.perf_counter() - t0)
            else:
                t0 = time.perf_counter()
                r1 = oracle_colid(A1, 40, 1e-6, seed=42)
                ts.append(time.perf_counter() - t0)
        c1[label] = {"ms_median": statistics.median(ts) * 1e3, "ms_min": min(ts) * 1e3,
                     "threads": lim if lim else threads}
    c1.update({"k": int(r1.k), "rel_error": float(id_error(A1, r1.perm, r1.k, r1.X)),
               "config": "1000x1000, s_i = exp(-i/20), b = 40, tol 1e-6, median of 10"})
    return {
        "value": 1.0 / dt,
        "unit": "IDs/s",
        "cores": threads,
        "kind": "port",
        "sample": f"median of {len(times)} full IDs of the same {A_host.shape[0]}x{A_host.shape[1]} "
                  f"matrix to tol {args.tol:g} (b={args.b}), {dt:.2f} s each; BLAS threads "
                  f"{threads}, host cores visible {cores} (affinity {affinity[0]}..{affinity[-1]})",
        "seconds": dt,
        "seconds_all": times,
        "nproc": os.cpu_count(),
        "affinity": [affinity[0], affinity[-1], cores],
        "k": int(r.k),
        "skeleton_equal_to_gpu": same,
        "config1": c1,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: the column ID runs on one GPU per call (asked for --gpus {args.gpus}); "
              "there is no multi-GPU form to measure", file=sys.stderr)
        sys.exit(2)
    import torch

    from robust_blockwise_random_pivoting_amd import Handle, rbrp_colid
    from robust_blockwise_random_pivoting_amd.matrices import exp_spectrum, spectrum_matrix_torch

    torch.cuda.set_device(0)
    m, n = args.m, args.n
    # the 50000 x 4000 matrix serves the main and mixed legs and the probes only
    need_a = (not args.no_main or not args.no_mixed or args.traffic_probe
              or bool(args.save_matrix))
    if not need_a:
        A = None
    elif args.load_matrix:
        At = np.load(args.load_matrix, mmap_mode="r")            # (n, m) row-major == A column-major
        A = torch.from_numpy(np.ascontiguousarray(At)).to("cuda").t()
        del At
    else:
        A = spectrum_matrix_torch(m, n, exp_spectrum(n, args.rate), seed=args.matrix_seed,
                                  device="cuda")
    torch.cuda.synchronize()
    if A is not None:
        print(f"[bench] matrix {m}x{n} ready", file=sys.stderr, flush=True)
    if args.save_matrix:
        np.save(args.save_matrix, A.t().contiguous().cpu().numpy())
        return
    if args.traffic_probe:
        # calibration: a plain device copy of exactly 2^30 bytes read and written
        src = torch.ones(1 << 27, dtype=torch.float64, device="cuda")
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst
        with Handle() as hp:
            if args.probe_what == "cpqr":
                from robust_blockwise_random_pivoting_amd import CPQR
                t_cpqr = CPQR(A, [min(m, n) // 2], handle=hp)[2]
                what = f"one CPQR ID t={t_cpqr:.3f}s"
            else:
                r = rbrp_colid(A, b=args.b, tol=args.tol, seed=args.seed, handle=hp)
                what = f"one ID k={r.k}"
        torch.cuda.synchronize()
        print(f"[bench] traffic probe: copy 1 GiB, {what}", file=sys.stderr, flush=True)
        return

    if args.no_main:
        out = {"mixed": None if args.no_mixed else mixed_leg(A, args),
               "batched": None if args.no_batched else batched_leg(args),
               "scale": None if args.no_scale else scale_leg(args),
               "ceilings": None if args.no_ceilings else ceilings_leg(),
               "sampler": None if args.no_sampler else sampler_leg(args)}
        print(json.dumps(out))
        return

    h = Handle()
    # the timed IDs record events around the projection and update launches
    # only (level 2); the per-phase breakdown comes from one more, untimed ID
    h.set_timing(1 if args.all_phase_timing else 2)
    res = None
    for w in range(args.warmup):
        res = rbrp_colid(A, b=args.b, tol=args.tol, seed=args.seed, handle=h)
        if args.profile_steps:
            print(f"[bench] warmup {w} k={res.k}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    h.reset_stats()
    t0 = time.perf_counter()
    for s in range(args.steps):
        res = rbrp_colid(A, b=args.b, tol=args.tol, seed=args.seed, handle=h)
        if args.profile_steps:
            print(f"[bench] step {s} k={res.k}", file=sys.stderr, flush=True)
    torch.cuda.synchronize()
    dt = time.perf_counter() - t0
    stats = h.stats()
    if args.all_phase_timing:
        pstats = stats
        phase_ms = {k: v["ms"] / args.steps for k, v in stats.items()}
    else:
        h.set_timing(1)
        h.reset_stats()
        rbrp_colid(A, b=args.b, tol=args.tol, seed=args.seed, handle=h)
        torch.cuda.synchronize()
        pstats = h.stats()
        phase_ms = {k: v["ms"] for k, v in pstats.items()}
        h.set_timing(0)

    # independent check of the returned ID on the device
    err = id_rel_error(A, res)

    kern = {}
    for name, bound in (("project", "mfma"), ("update", "hbm")):
        st = stats[name]
        if st["launches"] == 0 or st["ms"] <= 0:
            continue
        avg_ms = st["ms"] / st["launches"]
        if bound == "mfma":
            ach = st["flops"] / (st["ms"] * 1e-3) / 1e12
            kern[name] = dict(bound="mfma", achieved=ach, peak=PEAK_FP64_TFLOPS, unit="TFLOP/s",
                              frac=ach / PEAK_FP64_TFLOPS, avg_ms=avg_ms, launches=st["launches"],
                              flops_per_launch=st["flops"] / st["launches"])
        else:
            ach = st["bytes"] / (st["ms"] * 1e-3) / 1e9
            kern[name] = dict(bound="hbm", achieved=ach, peak=PEAK_HBM_GBS, unit="GB/s",
                              frac=ach / PEAK_HBM_GBS, avg_ms=avg_ms, launches=st["launches"],
                              bytes_per_launch=st["bytes"] / st["launches"])
    dom = max(kern, key=lambda k: stats[k]["ms"]) if kern else None
    # the initial column-norm pass (a read of A: e*m*n bytes, 8*n written) and
    # the per-block sampler (weights read, CDF written: 16 bytes per candidate),
    # timed by the phase events of the breakdown ID
    for name in ("norms", "sample"):
        st = pstats.get(name)
        if not st or st["launches"] == 0 or st["ms"] <= 0:
            continue
        ach = st["bytes"] / (st["ms"] * 1e-3) / 1e9
        kern[name] = dict(bound="hbm", achieved=ach, peak=PEAK_HBM_GBS, unit="GB/s",
                          frac=ach / PEAK_HBM_GBS, avg_ms=st["ms"] / st["launches"],
                          launches=st["launches"], bytes_per_launch=st["bytes"] / st["launches"],
                          timing="phase events, breakdown ID")
    roof = None
    if dom:
        d = kern[dom]
        roof = {"kernel": {"project": "k_gemm_tn64", "update": "k_gemm_upd64"}[dom], "bound": d["bound"],
                "achieved": d["achieved"], "peak": d["peak"], "unit": d["unit"], "frac": d["frac"],
                "traffic": measured_traffic(args) if dom == "update" else None,
                "algorithmic_bytes_per_launch": d.get("bytes_per_launch"),
                "avg_ms": d["avg_ms"], "launches": d["launches"]}

    # comparison baselines at the rank RBRP found: Gaussian
    # sketch of 2k rows, oversampling 2, interpolation as gpu/interp.cu
    comps = {}
    if not args.no_comparators:
        from robust_blockwise_random_pivoting_amd import RandCPQR, RandCPQR_OS, RandLUPP, RandLUPP_OS
        for name, fn in (("randcpqr_os", RandCPQR_OS), ("randlupp_os", RandLUPP_OS),
                         ("randcpqr", RandCPQR), ("randlupp", RandLUPP)):
            fn(A, res.k, seed=args.seed, handle=h)            # warm-up (workspace growth)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            sk, rd, T = fn(A, res.k, seed=args.seed, handle=h)
            torch.cuda.synchronize()
            dtc = time.perf_counter() - t1
            sk = torch.as_tensor(sk, device="cuda").long()
            rd = torch.as_tensor(rd, device="cuda").long()
            E = A[:, rd] - A[:, sk] @ T
            comps[name] = {"ms": dtc * 1e3, "rank": res.k,
                           "rel_error": (torch.linalg.norm(E) / torch.linalg.norm(A)).item()}
            del E
            print(f"[bench] {name}: {dtc * 1e3:.1f} ms", file=sys.stderr, flush=True)

    # full-matrix CPQR ID (gpu/id_cpqr.cu:30-172) at ranks of the RBRP trace,
    # as the reference's driver compares them (gpu/driver.cpp:197-205)
    cpqr = None
    if not args.no_cpqr:
        from robust_blockwise_random_pivoting_amd import CPQR
        nb = len(res.ranks)
        idx = sorted({max(nb // 4 - 1, 0), max(nb // 2 - 1, 0), max(3 * nb // 4 - 1, 0), nb - 1})
        ranks = [int(res.ranks[t]) for t in idx]
        print(f"[bench] cpqr at ranks {ranks} ...", file=sys.stderr, flush=True)
        e_q, t_q, t_full, fl_q, _ = CPQR(A, ranks, handle=h)
        q = min(m, n)
        # blocked pivoted QR: step j reads the (m-j) x (n-j-1) trailing block once (the F
        # column), each 32-column panel reads and writes the trailing block at its end
        qbytes = sum(8.0 * (m - j) * (n - j - 1) for j in range(q)) + \
            sum(16.0 * (m - r0) * (n - r0) for r0 in range(32, q, 32))
        cpqr = {"m": m, "n": n, "t_cpqr_ms": t_full * 1e3, "ranks": ranks,
                "rel_error": [float(x) ** 0.5 for x in e_q],
                "t_ranks_ms": [x * 1e3 for x in t_q], "flops": fl_q,
                "rbrp_rel_error": [float(res.error[t]) ** 0.5 for t in idx],
                "rbrp_t_ranks_ms": [float(res.t_ranks[t]) * 1e3 for t in idx],
                "qrcp_GBs": qbytes / t_full / 1e9, "qrcp_hbm_frac": qbytes / t_full / 1e9 / PEAK_HBM_GBS,
                "qrcp_bytes": qbytes}
        print(f"[bench] cpqr: {t_full * 1e3:.1f} ms", file=sys.stderr, flush=True)

    mixed = None if args.no_mixed else mixed_leg(A, args)
    batched = None if args.no_batched else batched_leg(args)
    scale = None if args.no_scale else scale_leg(args)
    ceilings = None if args.no_ceilings else ceilings_leg()
    sampler = None if args.no_sampler else sampler_leg(args)

    cpu = None
    if not args.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        A_host = A.cpu().numpy()
        cpu = cpu_baseline(A_host, args, res.k, res.perm.cpu().numpy())

    ms = dt / args.steps * 1e3
    out = {
        "metric": METRIC,
        "value": args.steps / dt,
        "unit": "IDs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded Haar U, V; s_i = exp(-i/100); generated on device)",
        "config": {"workload": f"FP64 RBRP column ID to tol, {m}x{n} exp-decay, b={args.b}, "
                               f"tol={args.tol:g}", "m": m, "n": n, "b": args.b, "tol": args.tol,
                   "parallelism": "single GPU"},
        "k": res.k,
        "blocks": len(res.ranks),
        "rel_error": err,
        "roofline": roof,
        "kernels": kern,
        "phase_ms_per_step": phase_ms,
        "comparators": comps,
        "cpqr": cpqr,
        "mixed": mixed,
        "batched": batched,
        "scale": scale,
        "ceilings": ceilings,
        "sampler": sampler,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
