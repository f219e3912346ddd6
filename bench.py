"""Benchmark: SGEMM 8192^3 (column-major, alpha 1, beta 0, uniform[-1,1] from
seed 0) on one MI355X, plain FP32-MFMA path vs 3-term split-BF16 MFMA path.

One step = one iclblasSgemm call (device-pointer mode, inputs resident in HBM)
on the handle's stream; K steps are timed with HIP events recorded on that
stream. Prints ONE JSON line:
  value        split-BF16 SGEMM throughput, useful TFLOP/s = 2 m n k / t
  paths        both paths: TFLOP/s, fraction of their MFMA peak, max error vs FP64
  roofline     the split-BF16 kernel: executed BF16 TFLOP/s (3 products) vs 2516.6
  roofline_fp32  the FP32-MFMA kernel vs 157.3
  cpu_baseline the reference's CPU GEMM (restated cpu_gemm, double accumulation,
               naive loop order) on the host cores, on a bounded column slab
Errors are measured against an FP64 product computed with torch on the GPU over
the whole 8192 x 8192 result: max |C - C64|, max relative error and the max of
|C - C64|_ij / (k 2^-24 (|A||B|)_ij) (must be <= 4).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

PEAK_FP32_MFMA = 157.3   # TFLOP/s, MI355X dense FP32 matrix
PEAK_BF16_MFMA = 2516.6  # TFLOP/s, MI355X dense BF16 matrix (no sparsity)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=100)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample duration")
    ap.add_argument("--paths", default="fp32,bf16x3")
    ap.add_argument("--impl", default="", help="comma list, one implementation name per path (empty = automatic)")
    ap.add_argument("--trans", default="NN", help="NN, NT, TN or TT for the timed op")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 error check (profiling runs)")
    ap.add_argument("--no-extra", action="store_true", help="skip the CGEMM / STRSM rows (same as --rows sgemm)")
    ap.add_argument("--rows", default="sgemm,sgemm_more,cgemm,strsm,ctrsm,chain,level3,gemv,host",
                    help="comma list of workloads to run: sgemm (the headline line), sgemm_more (the other SGEMM "
                         "configurations), cgemm, strsm, ctrsm, chain (the differentiable GEMM+TRSM chain), "
                         "level3 (SYRK, SYR2K, SYMM, TRMM, HERK, HEMM, complex TRMM), gemv (SGEMV, CGEMV), "
                         "host (host-pointer calls, PCIe included)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB device-to-device copy kernel first (scales HBM byte counters)")
    return ap.parse_args()


def _cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_threads() -> int:
    """Threads for the all-core leg: the CPUs this process may run on, capped by
    OMP_NUM_THREADS where the host sets it (the CPU share granted to this job)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    cap = os.environ.get("OMP_NUM_THREADS", "")
    if cap.isdigit() and int(cap) > 0:
        n = min(n, int(cap))
    return max(1, n)


def _slab_rate(run, m: int, n: int, k: int, step: int, target_s: float):
    """Times run(cols) on growing column slabs until one takes target_s; returns
    (TFLOP/s, cols, seconds)."""
    cols = step
    while True:
        dt = run(cols)
        if dt >= target_s or cols >= n:
            break
        cols = min(n, max(cols + step, int(cols * min(8.0, 1.2 * target_s / max(dt, 1e-3)))))
        cols = max(step, (cols // step) * step)
    return 2.0 * m * cols * k / dt / 1e12, cols, dt


def cpu_baseline(A: np.ndarray, B: np.ndarray, n: int, target_s: float) -> dict:
    """The reference's CPU GEMM on the host cores, on column slabs of the same
    workload, extrapolated linearly in n:
      - all cores: the restated cpu_gemm (oracle.sgemm_mt: the reference's loop
        order and double accumulator, output columns split over threads);
      - one thread: the reference's own cpu_gemm (tests/blas_funcs/test_gemm.hpp:83-201,
        compiled from its headers into oracle/_ref) where that build is present,
        else the restatement on one thread. The reference runs it single-threaded."""
    import oracle
    m = k = A.shape[0]
    threads = _cpu_threads()

    def run_mt(cols):
        C = np.zeros((m, cols), np.float32, order="F")
        Bs = np.asfortranarray(B[:, :cols])
        t0 = time.perf_counter()
        oracle.sgemm_mt(threads, 0, 0, m, cols, k, 1.0, A, Bs, 0.0, C)
        return time.perf_counter() - t0

    use_ref = oracle.ref_available()

    def run_1(cols):
        C = np.zeros((m, cols), np.float32, order="F")
        Bs = np.asfortranarray(B[:, :cols])
        t0 = time.perf_counter()
        if use_ref:
            oracle.sgemm(0, 0, m, cols, k, 1.0, A, Bs, 0.0, C, reference=True)
        else:
            oracle.sgemm_mt(1, 0, 0, m, cols, k, 1.0, A, Bs, 0.0, C)
        return time.perf_counter() - t0

    v_mt, c_mt, t_mt = _slab_rate(run_mt, m, n, k, threads, target_s)
    v_1, c_1, t_1 = _slab_rate(run_1, m, n, k, 1, max(2.0, target_s / 2))
    model = _cpu_model()
    return {"value": v_mt, "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "cpu_model": model, "host_cpus": os.cpu_count(),
            "affinity_cpus": len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else None,
            "sample": f"SGEMM NN {m}x{c_mt}x{k} column slab of the {m}x{n}x{k} workload "
                      f"({c_mt} of {n} output columns, {t_mt:.1f} s), naive cpu_gemm loop order with a double "
                      f"accumulator, columns split over {threads} threads on {model}; rate extrapolates linearly "
                      f"in n",
            "single_thread": {"value": v_1, "unit": "TFLOP/s", "cores": 1,
                              "kind": "reference" if use_ref else "port",
                              "sample": f"SGEMM NN {m}x{c_1}x{k} column slab ({c_1} of {n} columns, {t_1:.1f} s), "
                                        + ("the reference's own cpu_gemm compiled from its headers"
                                           if use_ref else "the restated cpu_gemm on one thread")}}


def _time(stream, fn, steps: int, warmup: int, settle_s: float = 0.25, warm_fn=None) -> float:
    """Mean ms per call over `steps` calls of fn, after `warmup` calls of warm_fn
    (default fn) and then as many more as keep the GPU busy for `settle_s`: an
    idle GPU's clock ramps up over the first tenths of a second of load, so a row
    timed right after host-side set-up would otherwise run its first calls at a
    lower clock."""
    import torch
    warm = warm_fn or fn
    t0 = time.perf_counter()
    n = 0
    while n < warmup or time.perf_counter() - t0 < settle_s:
        warm()
        n += 1
        if n % 4 == 0:
            torch.cuda.synchronize()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    for _ in range(steps):
        fn()
    e1.record(stream)
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps


def _solve_timer(stream, solve, rhs, steps: int) -> float:
    """Mean ms of an in-place solve, solve(X), over `steps` calls, each on its own
    copy of the right-hand side rhs made before the timed region (so every timed
    call solves the same data and no copy is timed); the warm-up solves a scratch
    copy restored before each call."""
    pool = [rhs.clone() for _ in range(steps)]
    it = iter(pool)
    scratch = rhs.clone()

    def warm():
        scratch.copy_(rhs)
        solve(scratch)
    ms = _time(stream, lambda: solve(next(it)), steps, 1, warm_fn=warm)
    del pool, scratch
    return ms


def sgemm_rows(h, ib, stream, steps: int) -> dict:
    """The other SGEMM configurations: 8192^3 NT / TN / TT, NN m=4096 n=8192
    k=2048 with lda=ldb=ldc=4160, and NN 256^3 with beta = 0.5, each on both
    paths, with the error against an FP64 product computed with torch on the GPU."""
    import torch

    from clgpu_amd import datagen
    dev = torch.device("cuda:0")
    out = {}

    def run(name, ta, tb, m, n, k, lda, ldb, ldc, beta):
        # column-major (ld x cols) storage as torch (cols, ld)
        acols, bcols = (k if ta == 0 else m), (n if tb == 0 else k)
        A = torch.from_numpy(np.ascontiguousarray(datagen.matrix(lda, acols, name="A", seed=1).T)).to(dev)
        B = torch.from_numpy(np.ascontiguousarray(datagen.matrix(ldb, bcols, name="B", seed=1).T)).to(dev)
        C0 = torch.from_numpy(np.ascontiguousarray(datagen.matrix(ldc, n, name="C", seed=1).T)).to(dev)
        # storage X_t[c, r] = X(r, c); op(A) as (m, k), op(B) as (k, n)
        opA = A.double()[:k, :m].t() if ta == 0 else A.double()[:m, :k]
        opB = B.double()[:n, :k].t() if tb == 0 else B.double()[:k, :n]
        ref = (opA @ opB + beta * C0.double()[:, :m].t()).t()  # storage layout (n, m)
        bound = (k * 2.0 ** -24 * (opA.abs() @ opB.abs() + abs(beta) * C0.double()[:, :m].t().abs())).t()
        res = {}
        for path, mode in (("fp32", ib.MATH_FP32), ("bf16x3", ib.MATH_SPLIT_BF16X3)):
            h.set_math_mode(mode)
            h.set_implementation(None)
            C = C0.clone()
            ib.check(ib.iclblasSgemm(h.value, ta, tb, m, n, k, 1.0, A.data_ptr(), lda, B.data_ptr(), ldb, beta,
                                     C.data_ptr(), ldc), "iclblasSgemm")
            torch.cuda.synchronize()
            err = (C.double()[:, :m] - ref).abs()
            pad_ok = bool(torch.equal(C[:, m:], C0[:, m:]))
            Cw = C0.clone()
            ms = _time(stream, lambda: ib.check(ib.iclblasSgemm(h.value, ta, tb, m, n, k, 1.0, A.data_ptr(), lda,
                                                                B.data_ptr(), ldb, beta, Cw.data_ptr(), ldc)),
                       steps, 3)
            tf = 2.0 * m * n * k / (ms * 1e-3) / 1e12
            res[path] = {"implementation": h.last_implementation(), "ms": ms, "tflops": tf,
                         "frac_of_fp32_mfma_peak": tf / PEAK_FP32_MFMA,
                         "max_abs_err": float(err.max()),
                         "max_err_over_bound_unit": float((err / bound.clamp_min(1e-300)).max()),
                         "padding_untouched": pad_ok}
            del C, Cw, err
        out[name] = res
        del A, B, C0, opA, opB, ref, bound
        torch.cuda.empty_cache()

    for tr in ("NT", "TN", "TT"):
        run(f"sgemm_8192_{tr}", int(tr[0] == "T"), int(tr[1] == "T"), 8192, 8192, 8192, 8192, 8192, 8192, 0.0)
    run("sgemm_4096x8192x2048_ld4160_NN", 0, 0, 4096, 8192, 2048, 4160, 4160, 4160, 0.0)
    run("sgemm_256_NN_beta0.5", 0, 0, 256, 256, 256, 256, 256, 256, 0.5)
    h.set_math_mode(ib.MATH_FP32)
    return out


def ctrsm_rows(h, ib, stream, steps: int) -> dict:
    """CTRSM left/lower/{N, C}/non-unit 4096 x 1024 (complex), FP32 path, the
    default panel solve (and op N with the recursive solve forced); forward error against a complex FP64 solve computed with torch on the GPU.
    Useful flops 4 m^2 n (one complex multiply-add = 8 real flops, m^2 n / 2 of them)."""
    import torch

    from clgpu_amd import datagen
    dev = torch.device("cuda:0")
    h.set_math_mode(ib.MATH_FP32)
    h.set_implementation(None)
    m, nr = 4096, 1024
    T = datagen.complex_triangular(m, lower=True, seed=0, poison=True)
    Bm = datagen.complex_matrix(m, nr, name="B", seed=0)
    tT = torch.from_numpy(np.ascontiguousarray(T.T)).to(dev)
    tB = torch.from_numpy(np.ascontiguousarray(Bm.T)).to(dev)
    L64 = torch.nan_to_num(tT.to(torch.complex128).t()).tril()
    out = {}
    for label, trans, impl in (("NN", 0, None), ("CN", 2, None), ("NN_recursive", 0, "ctrsm_recursive_inv64")):
        h.set_implementation(impl)
        opL = L64 if trans == 0 else L64.conj().t()
        X64 = torch.linalg.solve_triangular(opL, tB.to(torch.complex128).t(), upper=trans != 0)
        X = tB.clone()
        ib.check(ib.iclblasCtrsm(h.value, 0, 1, trans, 0, m, nr, 1.0, tT.data_ptr(), m, X.data_ptr(), m))
        torch.cuda.synchronize()
        fe = float((X.to(torch.complex128).t() - X64).norm() / X64.norm())
        ms = _solve_timer(stream, lambda Xw: ib.check(ib.iclblasCtrsm(h.value, 0, 1, trans, 0, m, nr, 1.0,
                                                                      tT.data_ptr(), m, Xw.data_ptr(), m)),
                          tB, max(2, steps // 4))
        tf = 4.0 * m * m * nr / (ms * 1e-3) / 1e12
        out[label] = {"implementation": h.last_implementation(), "ms": ms, "tflops": tf,
                      "frac_of_fp32_mfma_peak": tf / PEAK_FP32_MFMA, "forward_rel_err_fro": fe}
        del X, X64, opL
    h.set_implementation(None)
    del tT, tB, L64
    torch.cuda.empty_cache()
    return out


def level3_rows(h, ib, stream, steps: int) -> dict:
    """The GEMM-based level-3 routines on device-resident uniform[-1, 1] data
    (torch's generator, seed 0), on the FP32 path and in the split-BF16 math
    mode: SSYRK, SSYR2K, SSYMM, STRMM at 8192, CHERK, CHEMM, CTRMM at 4096.
    Useful flops (real flops; a complex multiply-add counts 8): SYRK n^2 k,
    SYR2K 2 n^2 k, SYMM 2 m^2 n, TRMM m^2 n, HERK 4 n^2 k, HEMM 8 m^2 n, complex
    TRMM 4 m n^2 (right side). Each result is compared with an FP64 product on
    the GPU (normwise relative error over the written part)."""
    import torch
    dev = torch.device("cuda:0")
    h.set_implementation(None)
    g = torch.Generator(device=dev)
    g.manual_seed(0)

    def rnd(rows, cols, cplx=False):
        """(cols, rows) tensor = column-major rows x cols matrix, ld = rows."""
        x = torch.rand((cols, rows), generator=g, device=dev) * 2 - 1
        if cplx:
            x = torch.complex(x, torch.rand((cols, rows), generator=g, device=dev) * 2 - 1)
        return x

    def f64(x):  # the matrix (rows x cols) in float64 / complex128
        return x.t().to(torch.complex128 if x.is_complex() else torch.float64)

    def rel(got, exp, mask=None):
        d = f64(got) - exp
        if mask is not None:
            d, exp = d * mask, exp * mask
        return float(d.norm() / exp.norm())

    def row(call, flops, err, reset=None):
        """Both math modes: time the call, then (after reset()) run it once more
        and take err() of the result."""
        res = {}
        for path, mode in (("fp32", ib.MATH_FP32), ("bf16x3", ib.MATH_SPLIT_BF16X3)):
            h.set_math_mode(mode)
            call()
            torch.cuda.synchronize()
            ms = _time(stream, call, max(2, steps // 4), 1)
            tf = flops / (ms * 1e-3) / 1e12
            r = {"implementation": h.last_implementation(), "ms": ms, "tflops": tf,
                 "frac_of_fp32_mfma_peak": tf / PEAK_FP32_MFMA}
            if reset:
                reset()
            call()
            torch.cuda.synchronize()
            r["normwise_rel_err"] = err()
            res[path] = r
        h.set_math_mode(ib.MATH_FP32)
        return res

    out = {}
    n = 8192
    # SSYRK lower, op N: C = A A^T + 0.5 C
    A, C = rnd(n, n), rnd(n, n)
    C0 = C.clone()
    low = torch.ones((n, n), device=dev, dtype=torch.float64).tril()
    out["ssyrk_8192_k8192_LN"] = row(
        lambda: ib.check(ib.iclblasSsyrk(h.value, 1, 0, n, n, 1.0, A.data_ptr(), n, 0.5, C.data_ptr(), n)),
        float(n) * n * n, lambda: rel(C, (f64(A) @ f64(A).t() + 0.5 * f64(C0)) * low, low), lambda: C.copy_(C0))
    # SSYR2K upper, op T, k = 4096
    k = 4096
    A2, B2 = rnd(k, n), rnd(k, n)
    C.copy_(C0)
    up = low.t()
    P = f64(A2).t() @ f64(B2)
    out["ssyr2k_8192_k4096_UT"] = row(
        lambda: ib.check(ib.iclblasSsyr2k(h.value, 0, 1, n, k, 1.0, A2.data_ptr(), k, B2.data_ptr(), k, 0.0,
                                          C.data_ptr(), n)),
        2.0 * n * n * k, lambda: rel(C, (P + P.t()) * up, up))
    del A2, B2, P
    # SSYMM left, lower
    Bm = rnd(n, n)
    S = f64(A).tril()
    S = S + S.tril(-1).t()
    out["ssymm_8192_LL"] = row(
        lambda: ib.check(ib.iclblasSsymm(h.value, 0, 1, n, n, 1.0, A.data_ptr(), n, Bm.data_ptr(), n, 0.0,
                                         C.data_ptr(), n)),
        2.0 * n * n * n, lambda: rel(C, S @ f64(Bm)))
    # STRMM left, lower, op N, non-unit (out of place)
    out["strmm_8192_LLN"] = row(
        lambda: ib.check(ib.iclblasStrmm(h.value, 0, 1, 0, 0, n, n, 1.0, A.data_ptr(), n, Bm.data_ptr(), n,
                                         C.data_ptr(), n)),
        float(n) * n * n, lambda: rel(C, f64(A).tril() @ f64(Bm)))
    del A, C, C0, Bm, S, low, up
    torch.cuda.empty_cache()

    n = 4096
    A, C, Bm = rnd(n, n, True), rnd(n, n, True), rnd(n, n, True)
    # CHERK upper, op N, alpha 1, beta 0
    up = torch.ones((n, n), device=dev, dtype=torch.float64).triu().to(torch.complex128)
    out["cherk_4096_k4096_UN"] = row(
        lambda: ib.check(ib.iclblasCherk(h.value, 0, 0, n, n, 1.0, A.data_ptr(), n, 0.0, C.data_ptr(), n)),
        4.0 * n * n * n, lambda: rel(C, (f64(A) @ f64(A).conj().t()) * up, up))
    # CHEMM left, upper
    Hm = f64(A).triu()
    Hm = Hm + Hm.triu(1).conj().t()
    Hm.diagonal().imag.zero_()
    out["chemm_4096_LU"] = row(
        lambda: ib.check(ib.iclblasChemm(h.value, 0, 0, n, n, 1.0, A.data_ptr(), n, Bm.data_ptr(), n, 0.0,
                                         C.data_ptr(), n)),
        8.0 * n * n * n, lambda: rel(C, Hm @ f64(Bm)))
    # CTRMM right, upper, op C, non-unit
    out["ctrmm_4096_RUC"] = row(
        lambda: ib.check(ib.iclblasCtrmm(h.value, 1, 0, 2, 0, n, n, 1.0, A.data_ptr(), n, Bm.data_ptr(), n,
                                         C.data_ptr(), n)),
        4.0 * n * n * n, lambda: rel(C, f64(Bm) @ f64(A).triu().conj().t()))
    del A, C, Bm, Hm, up
    torch.cuda.empty_cache()
    return out


def gemv_rows(h, ib, stream, steps: int) -> dict:
    """SGEMV 16384 x 16384 (op N and T) and CGEMV 8192 x 8192 (op N and C) on
    device-resident uniform[-1, 1] data, beta = 1. Algorithmic HBM bytes per
    call: A once (m n elem), x once, y read and written; reported as GB/s
    against the 8 TB/s HBM peak, with the error against an FP64 product."""
    import torch
    dev = torch.device("cuda:0")
    g = torch.Generator(device=dev)
    g.manual_seed(0)
    out = {}
    for label, cplx, n, trans in (("sgemv_16384_N", False, 16384, 0), ("sgemv_16384_T", False, 16384, 1),
                                  ("cgemv_8192_N", True, 8192, 0), ("cgemv_8192_C", True, 8192, 2)):
        def rnd(*shape):
            x = torch.rand(shape, generator=g, device=dev) * 2 - 1
            return torch.complex(x, torch.rand(shape, generator=g, device=dev) * 2 - 1) if cplx else x
        A, x, y = rnd(n, n), rnd(n), rnd(n)
        y0 = y.clone()
        fn = ib.iclblasCgemv if cplx else ib.iclblasSgemv
        call = lambda: ib.check(fn(h.value, trans, n, n, 1.0, A.data_ptr(), n, x.data_ptr(), 1, 1.0, y.data_ptr(), 1))
        call()
        torch.cuda.synchronize()
        ms = _time(stream, call, max(10, steps * 2), 2)
        es = 8 if cplx else 4
        nbytes = float(n) * n * es + 3.0 * n * es
        y.copy_(y0)
        call()
        torch.cuda.synchronize()
        M = A.t().to(torch.complex128 if cplx else torch.float64)
        op = M if trans == 0 else (M.conj().t() if trans == 2 else M.t())
        ref = op @ x.to(M.dtype) + y0.to(M.dtype)
        out[label] = {"implementation": h.last_implementation(), "ms": ms, "GBps": nbytes / (ms * 1e-3) / 1e9,
                      "frac_of_hbm_peak": nbytes / (ms * 1e-3) / 8e12,
                      "normwise_rel_err": float((y.to(M.dtype) - ref).norm() / ref.norm())}
        del A, x, y, y0, M, op, ref
        torch.cuda.empty_cache()
    return out


def host_rows(ib, steps: int) -> dict:
    """The reference's calling convention: host pointers, synchronous calls, so
    every call moves its matrices over PCIe (through the handle's pinned
    buffers). Wall-clock time per call on numpy arrays: SGEMM 8192^3 NN (FP32 and
    split-BF16; A and B in, C out, beta = 0 so C is not sent), STRSM 4096 x 1024
    (A and B in, B out) and SGEMV 16384^2 (A in). These PCIe-inclusive rates are
    reported beside the device-resident ones, never as the headline value."""
    import time
    rng = np.random.default_rng(0)
    out = {}

    def wall(fn, k):
        fn()
        t0 = time.perf_counter()
        for _ in range(k):
            fn()
        return (time.perf_counter() - t0) / k

    k = max(2, steps // 4)
    with ib.Handle() as h:
        n = 8192
        A = np.asfortranarray(rng.uniform(-1, 1, (n, n)).astype(np.float32))
        B = np.asfortranarray(rng.uniform(-1, 1, (n, n)).astype(np.float32))
        C = np.asfortranarray(np.zeros((n, n), np.float32))
        moved = 3.0 * n * n * 4
        for label, mode in (("sgemm_8192_fp32", ib.MATH_FP32), ("sgemm_8192_split", ib.MATH_SPLIT_BF16X3)):
            h.set_math_mode(mode)
            t = wall(lambda: ib.check(ib.iclblasSgemm(h.value, 0, 0, n, n, n, 1.0, A, n, B, n, 0.0, C, n)), k)
            out[label] = {"ms": t * 1e3, "tflops": 2.0 * n ** 3 / t / 1e12, "pcie_bytes": moved,
                          "pcie_GBps_if_all_transfer": moved / t / 1e9, "implementation": h.last_implementation()}
        h.set_math_mode(ib.MATH_FP32)
        m, nr = 4096, 1024
        T = np.asfortranarray(np.tril(rng.uniform(-1, 1, (m, m)) / 64 + np.diag(rng.uniform(1, 2, m))).astype(np.float32))
        Bt = np.asfortranarray(rng.uniform(-1, 1, (m, nr)).astype(np.float32))
        X = Bt.copy(order="F")

        def trsm():
            X[:] = Bt
            ib.check(ib.iclblasStrsm(h.value, 0, 1, 0, 0, m, nr, 1.0, T, m, X, m))
        t = wall(trsm, k)
        moved = (m * m + 2.0 * m * nr) * 4
        out["strsm_4096x1024"] = {"ms": t * 1e3, "tflops": float(m) * m * nr / t / 1e12, "pcie_bytes": moved,
                                  "pcie_GBps_if_all_transfer": moved / t / 1e9}
        del T, Bt, X
        n = 16384
        Ag = np.asfortranarray(rng.uniform(-1, 1, (n, n)).astype(np.float32))
        x = rng.uniform(-1, 1, n).astype(np.float32)
        y = np.zeros(n, np.float32)
        t = wall(lambda: ib.check(ib.iclblasSgemv(h.value, 0, n, n, 1.0, Ag, n, x, 1, 0.0, y, 1)), k)
        moved = (n * n + 2.0 * n) * 4
        out["sgemv_16384"] = {"ms": t * 1e3, "pcie_bytes": moved, "pcie_GBps_if_all_transfer": moved / t / 1e9}
    return out


def chain_rows(h, ib, stream, steps: int) -> dict:
    """The differentiable chain X = A^-1 (alpha G H), loss = sum(W * X) with A
    2048^2 lower triangular, G 2048^2, H 2048 x 512, alpha = 0.5: one step is the
    forward pass (SGEMM, STRSM) and the backward pass (STRSM adjoint: dY and the
    masked dA; SGEMM adjoint: dG, dH) through the C ABI on device pointers.
    Useful flops: forward 2 n^2 r + n^2 r, backward n^2 r + 3 * 2 n^2 r."""
    import torch

    from clgpu_amd import datagen
    dev = torch.device("cuda:0")
    h.set_math_mode(ib.MATH_FP32)
    h.set_implementation(None)
    n, r, alpha = 2048, 512, 0.5

    def t(x):  # column-major host matrix -> torch (cols, ld) on the GPU
        return torch.from_numpy(np.ascontiguousarray(np.asarray(x, np.float32).T)).to(dev)

    A = t(datagen.triangular(n, lower=True, seed=0, poison=False))
    G = t(datagen.matrix(n, n, name="G", seed=0))
    H = t(datagen.matrix(n, r, name="H", seed=0))
    W = t(datagen.normal(n * r, seed=0, stream=5).reshape(r, n).T)
    X = torch.empty((r, n), device=dev)
    dY = torch.empty((r, n), device=dev)
    dA = torch.empty((n, n), device=dev)
    dG = torch.empty((n, n), device=dev)
    dH = torch.empty((r, n), device=dev)

    def step():
        ib.check(ib.iclblasSgemm(h.value, 0, 0, n, r, n, alpha, G.data_ptr(), n, H.data_ptr(), n, 0.0, X.data_ptr(),
                                 n), "iclblasSgemm")
        ib.check(ib.iclblasStrsm(h.value, 0, 1, 0, 0, n, r, 1.0, A.data_ptr(), n, X.data_ptr(), n), "iclblasStrsm")
        ib.check(ib.iclblasStrsmBackward(h.value, 0, 1, 0, 0, n, r, 1.0, A.data_ptr(), n, X.data_ptr(), n,
                                         W.data_ptr(), n, dA.data_ptr(), n, dY.data_ptr(), n), "iclblasStrsmBackward")
        ib.check(ib.iclblasSgemmBackward(h.value, 0, 0, n, r, n, alpha, G.data_ptr(), n, H.data_ptr(), n,
                                         dY.data_ptr(), n, dG.data_ptr(), n, dH.data_ptr(), n), "iclblasSgemmBackward")

    step()
    torch.cuda.synchronize()
    # FP64 check of the loss and of dH (dH = alpha G^T A^-T W)
    L64 = A.double().t().tril()
    X64 = torch.linalg.solve_triangular(L64, alpha * (G.double().t() @ H.double().t()), upper=False)
    loss64 = float((W.double().t() * X64).sum())
    loss = float((W.double() * X.double()).sum())
    dY64 = torch.linalg.solve_triangular(L64.t(), W.double().t(), upper=True)
    dH64 = alpha * G.double() @ dY64  # (G^T)^T storage: G.t() is G, so G^T dY64 = G_storage @ dY64
    dH_err = float((dH.double().t() - dH64).norm() / dH64.norm())
    ms = _time(stream, step, max(2, steps // 2), 1)
    flops = 3.0 * n * n * r + 7.0 * n * n * r
    out = {"ms": ms, "tflops": flops / (ms * 1e-3) / 1e12, "loss_rel_err": abs(loss - loss64) / abs(loss64),
           "dH_rel_err": dH_err}
    del A, G, H, W, X, dY, dA, dG, dH, L64, X64, dY64, dH64
    torch.cuda.empty_cache()
    return out


def cgemm_rows(h, ib, stream, steps: int) -> dict:
    """CGEMM 4096^3 NN and CN (alpha = (1, 0.5), beta = (0.25, 0)) on both
    paths; errors against an FP64 complex product computed with torch on the GPU."""
    import torch

    from clgpu_amd import datagen
    dev = torch.device("cuda:0")
    n = 4096
    A = torch.from_numpy(np.ascontiguousarray(datagen.complex_matrix(n, n, name="A").T)).to(dev)
    B = torch.from_numpy(np.ascontiguousarray(datagen.complex_matrix(n, n, name="B").T)).to(dev)
    C0 = torch.from_numpy(np.ascontiguousarray(datagen.complex_matrix(n, n, name="C").T)).to(dev)
    alpha, beta = 1 + 0.5j, 0.25
    flops = 8.0 * n ** 3
    cg = {}
    for opname, ta in (("NN", 0), ("CN", 2)):
        A64 = A.to(torch.complex128)
        opAT = A64 if ta == 0 else A64.conj().t()  # storage of op(A)^T
        ref = alpha * (B.to(torch.complex128) @ opAT) + beta * C0.to(torch.complex128)
        absab = (B.to(torch.complex128).abs() @ opAT.abs())  # |z| bound, a little loose vs |re|+|im|
        res = {}
        for path, mode in (("fp32", ib.MATH_FP32), ("bf16x3", ib.MATH_SPLIT_BF16X3)):
            h.set_math_mode(mode)
            h.set_implementation(None)
            C = C0.clone()

            def step():
                ib.check(ib.iclblasCgemm(h.value, ta, 0, n, n, n, alpha, A.data_ptr(), n, B.data_ptr(), n, beta,
                                         C.data_ptr(), n), "iclblasCgemm")
            step()
            torch.cuda.synchronize()
            err = (C.to(torch.complex128) - ref).abs()
            ratio = float((err / (n * 2.0 ** -24 * (abs(alpha) * 2 * absab + abs(beta) * C0.abs().double()))).max())
            Cw = C0.clone()
            ms = _time(stream, lambda: ib.check(ib.iclblasCgemm(h.value, ta, 0, n, n, n, alpha, A.data_ptr(), n,
                                                                B.data_ptr(), n, beta, Cw.data_ptr(), n)), steps, 2)
            tf = flops / (ms * 1e-3) / 1e12
            res[path] = {"implementation": h.last_implementation(), "ms": ms, "tflops": tf,
                         "frac_of_fp32_mfma_peak": tf / PEAK_FP32_MFMA,
                         "executed_frac_of_mfma_peak": (tf / PEAK_FP32_MFMA) if path == "fp32"
                         else 3 * tf / PEAK_BF16_MFMA,
                         "max_abs_err": float(err.max()), "normwise_rel_err": float(err.norm() / ref.abs().norm()),
                         "max_err_over_bound_unit": ratio}
            del err, C, Cw
        cg[opname] = res
        del ref, absab, A64, opAT
    h.set_math_mode(ib.MATH_FP32)
    del A, B, C0
    torch.cuda.empty_cache()
    return cg


def strsm_rows(h, ib, stream, steps: int) -> dict:
    """STRSM left/lower/N/non-unit 4096 x 1024, plain and with at most 3
    refinement sweeps (they stop once a correction no longer changes X at FP32
    resolution; `sweeps_run` reports how many ran); forward error against an
    FP64 solve computed with torch on the GPU. Every timed call solves its own copy
    of the same right-hand side, made before the timed region."""
    import torch

    from clgpu_amd import datagen
    dev = torch.device("cuda:0")
    h.set_math_mode(ib.MATH_FP32)
    h.set_implementation(None)
    m, nr = 4096, 1024
    T = datagen.triangular(m, lower=True, seed=0, poison=True)
    Bm = datagen.matrix(m, nr, name="B", seed=0)
    tT = torch.from_numpy(np.ascontiguousarray(T.T)).to(dev)
    tB = torch.from_numpy(np.ascontiguousarray(Bm.T)).to(dev)
    L64 = torch.nan_to_num(tT.double().t()).tril()
    X64 = torch.linalg.solve_triangular(L64, tB.double().t(), upper=False)
    flops_t = float(m) * m * nr
    ts = {}
    for label, sweeps in (("plain", 0), ("refined_max_3_sweeps", 3)):
        X = tB.clone()
        ib.check(ib.iclblasStrsmRefined(h.value, 0, 1, 0, 0, m, nr, 1.0, tT.data_ptr(), m, X.data_ptr(), m, sweeps))
        torch.cuda.synchronize()
        fe = float((X.double().t() - X64).norm() / X64.norm())
        sweeps_run = h.last_refinement_sweeps()
        ms = _solve_timer(stream, lambda Xw: ib.check(ib.iclblasStrsmRefined(
            h.value, 0, 1, 0, 0, m, nr, 1.0, tT.data_ptr(), m, Xw.data_ptr(), m, sweeps)), tB, max(2, steps // 4))
        tf = flops_t / (ms * 1e-3) / 1e12
        ts[label] = {"implementation": h.last_implementation(), "ms": ms, "tflops": tf,
                     "frac_of_fp32_mfma_peak": tf / PEAK_FP32_MFMA, "forward_rel_err_fro": fe,
                     "sweeps_run": sweeps_run}
    del tT, tB, L64, X64, X
    torch.cuda.empty_cache()
    return ts


def sgemm_line(a, h, ib, stream, dev):
    """The headline SGEMM measurement: both paths, errors, roofline fields.
    Returns the JSON line and the host inputs (for the CPU baseline)."""
    import torch

    from clgpu_amd import datagen
    n = a.size
    m = k = n
    ta, tb = ({"N": 0, "T": 1}[a.trans[0]], {"N": 0, "T": 1}[a.trans[1]])
    print(f"[bench] generating {n}x{n} inputs", file=sys.stderr, flush=True)
    A = datagen.matrix(n, n, name="A", seed=0)
    B = datagen.matrix(n, n, name="B", seed=0)
    # column-major (ld x cols) <-> torch (cols, ld) row-major
    tA = torch.from_numpy(np.ascontiguousarray(A.T)).to(dev)
    tB = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev)
    tC = torch.empty((n, n), device=dev, dtype=torch.float32)

    C64 = absAB = None
    if not a.no_check:
        # FP64 reference of the stored result: storage(C) = C^T = op(B)^T op(A)^T
        print("[bench] FP64 reference", file=sys.stderr, flush=True)
        sA = tA.double()
        sB = tB.double()
        opAT = sA if ta == 0 else sA.t()  # storage(op(A)) transposed view
        opBT = sB if tb == 0 else sB.t()
        C64 = opBT @ opAT
        absAB = opBT.abs() @ opAT.abs()
        del sA, sB, opAT, opBT
        torch.cuda.synchronize()

    flops = 2.0 * m * n * k
    results = {}
    impls = a.impl.split(",") if a.impl else []
    for pi, path in enumerate(a.paths.split(",")):
        mode = ib.MATH_FP32 if path == "fp32" else ib.MATH_SPLIT_BF16X3
        h.set_math_mode(mode)
        h.set_implementation(impls[pi] if pi < len(impls) else None)

        def step():
            ib.check(ib.iclblasSgemm(h.value, ta, tb, m, n, k, 1.0, tA.data_ptr(), n, tB.data_ptr(), n, 0.0,
                                     tC.data_ptr(), n), "iclblasSgemm")

        for _ in range(a.warmup):
            step()
        torch.cuda.synchronize()
        # The library records HIP events around the main GEMM kernel of every call
        # (the handle's stream, the one the kernel runs on), over the same timed
        # region as the call. At 8192^2 the split path packs the first halves of
        # its operands outside the pair, then runs its product as three launches
        # (A0 x B0 beside the pack of the second halves on a second stream, then
        # A0 x B1 and A1 x B), all inside the pair.
        h.set_kernel_timing(True)
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(a.steps):
            step()
        stop.record(stream)
        torch.cuda.synchronize()
        kern_total, kern_launches = h.kernel_timing()
        h.set_kernel_timing(False)
        ms = start.elapsed_time(stop) / a.steps
        kern_ms = kern_total / kern_launches if kern_launches > 0 else None
        impl = h.last_implementation()
        tflops = flops / (ms * 1e-3) / 1e12
        peak = PEAK_FP32_MFMA if path == "fp32" else PEAK_BF16_MFMA
        per_launch = flops if path == "fp32" else 3.0 * flops  # executed MFMA flops of one launch
        executed = per_launch / (kern_ms * 1e-3) / 1e12 if kern_ms else None
        results[path] = {"implementation": impl, "ms": ms, "tflops": tflops,
                         "frac_of_fp32_mfma_peak": tflops / PEAK_FP32_MFMA,
                         "kernel_ms": kern_ms, "kernel_launches": kern_launches,
                         "executed_tflops": executed, "mfma_peak": peak,
                         "mfma_frac": executed / peak if executed else None}
        if C64 is not None:
            err = (tC.double() - C64).abs()
            ratio = float((err / (k * 2.0 ** -24 * absAB)).max())
            results[path].update({
                "max_abs_err": float(err.max()),
                "max_rel_err": float((err / C64.abs().clamp_min(1e-300)).max()),
                "normwise_rel_err": float(err.norm() / C64.norm()),
                "max_err_over_bound_unit": ratio,
                "within_bound": ratio <= 4.0})
            del err
        print(f"[bench] {path}: {ms:.3f} ms  {tflops:.1f} TFLOP/s  {results[path].get('max_err_over_bound_unit')}"
              f"  ({impl})", file=sys.stderr, flush=True)

    head = "bf16x3" if "bf16x3" in results else "fp32"
    hr = results[head]
    # HBM bytes per launch come from the rocprofv3 FETCH_SIZE / WRITE_SIZE passes
    # summarised in profiles/sgemm_8192_hbm_traffic.md and recorded per kernel in
    # profiles/hbm_traffic.json (a live run has no counters); the value is used
    # only when the record names this kernel variant and this workload.
    traffic_rec = {}
    tf = ROOT / "profiles" / "hbm_traffic.json"
    if tf.exists():
        traffic_rec = json.loads(tf.read_text())

    def traffic_of(impl):
        r = traffic_rec.get(impl)
        if isinstance(r, dict) and r.get("workload") == f"SGEMM {a.trans} {n}^3":
            return r.get("bytes_per_call")
        return None
    traffic = traffic_of(hr["implementation"])
    line = {
        "metric": "SGEMM TFLOP/s @8192: split-BF16 vs FP32-MFMA path, %peak, max err vs FP64",
        "value": hr["tflops"],
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": hr["ms"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 via bf16x3 split (f32 accumulate)" if head == "bf16x3" else "f32",
        "data": "synthetic: uniform[-1,1] SplitMix64 seed 0, column-major",
        "config": {"workload": f"SGEMM {a.trans} m=n=k={n} alpha=1 beta=0 lda=ldb=ldc={n}, device-resident",
                   "path": head},
        "paths": results,
        "roofline": {"bound": "mfma", "achieved": hr["executed_tflops"], "peak": hr["mfma_peak"],
                     "unit": "TFLOP/s", "frac": hr["mfma_frac"], "traffic": traffic,
                     "kernel": hr["implementation"], "kernel_ms": hr["kernel_ms"],
                     "note": "achieved = executed MFMA flops of the call's product (2mnk; 3 x 2mnk for "
                             "split-BF16) / mean duration of the HIP-event pair around it (same timed region): the "
                             "one FP32 kernel; for split-BF16 at 8192^2 the three partial-product launches, one of "
                             "them beside the side-stream pack of the operands' second halves (the first halves' "
                             "pack is outside the pair); traffic = HBM bytes of those launches per call from the "
                             "rocprofv3 FETCH_SIZE / WRITE_SIZE passes of profiles/sgemm_8192_hbm_traffic.md "
                             "(profiles/hbm_traffic.json)"},
    }
    if "fp32" in results:
        r = results["fp32"]
        line["roofline_fp32"] = {"bound": "mfma", "achieved": r["executed_tflops"], "peak": PEAK_FP32_MFMA,
                                 "unit": "TFLOP/s", "frac": r["mfma_frac"],
                                 "traffic": traffic_of(r["implementation"]),
                                 "kernel": r["implementation"], "kernel_ms": r["kernel_ms"]}
        if "bf16x3" in results:
            line["split_over_fp32"] = results["bf16x3"]["tflops"] / r["tflops"]
    del C64, absAB, tA, tB, tC
    torch.cuda.empty_cache()
    return line, A, B


# Pass criteria of the rows: the elementwise bound of the north-star contract for
# GEMM products (max |err| / (k 2^-24 (|A||B|)) <= 4), FP32 accuracy (2^-23) for the
# refined solve, and a normwise tolerance for the solves, level-3 and GEMV rows.
ROW_TOL = {"strsm_plain": 1e-5, "strsm_refined": 2.0 ** -23, "ctrsm": 1e-5, "level3": 1e-5, "gemv": 1e-5,
           "chain": 1e-4}


def row_checks(line: dict):
    """One (ok, text) summary per measured configuration and path."""
    out = []

    def fmt(name, r, err_key, tol):
        err = r.get(err_key)
        ok = err is None or (err == err and err <= tol)  # NaN fails
        rate = f"{r['tflops']:.1f} TFLOP/s" if "tflops" in r else (f"{r['GBps']:.0f} GB/s" if "GBps" in r else "")
        ms = f"{r['ms']:.3f} ms" if r.get("ms") is not None else ""
        e = f"{err_key}={err:.3g} (<= {tol:.3g})" if err is not None else "unchecked"
        out.append((ok, f"{'ok  ' if ok else 'FAIL'} {name}: {ms} {rate} {e} [{r.get('implementation', '')}]"))

    for path, r in line.get("paths", {}).items():
        fmt(f"sgemm_8192_{line['config'].get('workload', '').split()[1]}/{path}", r, "max_err_over_bound_unit", 4.0)
    other = line.get("other_rows", {})
    for name, rows in other.items():
        if name.startswith("sgemm_"):
            for path, r in rows.items():
                fmt(f"{name}/{path}", r, "max_err_over_bound_unit", 4.0)
        elif name == "cgemm_4096":
            for op, paths in rows.items():
                for path, r in paths.items():
                    fmt(f"cgemm_4096_{op}/{path}", r, "max_err_over_bound_unit", 4.0)
        elif name == "strsm_4096x1024":
            for label, r in rows.items():
                fmt(f"strsm_4096x1024/{label}", r, "forward_rel_err_fro",
                    ROW_TOL["strsm_plain" if label == "plain" else "strsm_refined"])
        elif name == "ctrsm_4096x1024":
            for label, r in rows.items():
                fmt(f"ctrsm_4096x1024/{label}", r, "forward_rel_err_fro", ROW_TOL["ctrsm"])
        elif name == "chain_2048x512":
            fmt("chain_2048x512/loss", rows, "loss_rel_err", ROW_TOL["chain"])
            fmt("chain_2048x512/dH", rows, "dH_rel_err", ROW_TOL["chain"])
        elif name == "level3":
            for label, paths in rows.items():
                for path, r in paths.items():
                    fmt(f"{label}/{path}", r, "normwise_rel_err", ROW_TOL["level3"])
        elif name == "gemv":
            for label, r in rows.items():
                fmt(label, r, "normwise_rel_err", ROW_TOL["gemv"])
        elif name == "host_pointer":
            for label, r in rows.items():
                fmt(f"host/{label}", r, "none", 0.0)
    return out


def device_info() -> dict:
    """The gfx950 agent's name, compute units and peak engine clock from rocminfo
    (a child process, run before this process touches the GPU), and the MFMA
    peaks they give: 256 FP32 and 4096 dense BF16 flops per clock per CU."""
    import re
    import subprocess
    try:
        out = subprocess.run(["rocminfo"], capture_output=True, text=True, timeout=60).stdout
    except (OSError, subprocess.SubprocessError):
        return {}
    for block in re.split(r"\n\s*Agent \d+", out)[1:]:
        if "gfx950" not in block:
            continue
        cus = re.search(r"Compute Unit:[ \t]+(\d+)", block)
        mhz = re.search(r"Max Clock Freq\. \(MHz\):[ \t]+(\d+)", block)
        name = re.search(r"Marketing Name:[ \t]*(.*)", block)
        if not (cus and mhz):
            continue
        c, f = int(cus.group(1)), int(mhz.group(1))
        return {"name": (name.group(1).strip() if name else "") or "gfx950", "arch": "gfx950", "compute_units": c,
                "max_engine_clock_mhz": f, "fp32_mfma_peak_tflops": c * 256 * f * 1e6 / 1e12,
                "bf16_mfma_peak_tflops": c * 4096 * f * 1e6 / 1e12}
    return {}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (this build has no multi-GPU path)",
              file=sys.stderr)
        sys.exit(2)
    rows = ["sgemm"] if a.no_extra else [r for r in a.rows.split(",") if r]
    unknown = set(rows) - {"sgemm", "sgemm_more", "cgemm", "strsm", "ctrsm", "chain", "level3", "gemv", "host"}
    if unknown:
        print(f"bench.py: unknown rows {sorted(unknown)}", file=sys.stderr)
        sys.exit(2)
    device = device_info()
    import torch

    from clgpu_amd import iclblas as ib

    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    stream = torch.cuda.current_stream()
    h = ib.Handle(memory_mode=ib.MEMORY_DEVICE, stream=stream.cuda_stream)

    if a.copy_calibration:
        # a plain copy of known size: 1 GiB read + 1 GiB written
        src = torch.empty(1 << 28, device=dev, dtype=torch.float32).fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        del src, dst

    A = B = None
    if "sgemm" in rows:
        line, A, B = sgemm_line(a, h, ib, stream, dev)
    else:
        # a run of the secondary rows only (profiling): no headline value
        line = {"metric": "SGEMM TFLOP/s @8192: split-BF16 vs FP32-MFMA path, %peak, max err vs FP64",
                "value": None, "unit": "TFLOP/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
                "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
                "dtype": "f32", "data": "synthetic", "config": {"workload": "secondary rows only: " + ",".join(rows)}}
    other = {}
    if "sgemm_more" in rows:
        print("[bench] other SGEMM rows", file=sys.stderr, flush=True)
        other.update(sgemm_rows(h, ib, stream, a.steps))
    if "cgemm" in rows:
        print("[bench] CGEMM rows", file=sys.stderr, flush=True)
        other["cgemm_4096"] = cgemm_rows(h, ib, stream, a.steps)
    if "strsm" in rows:
        print("[bench] STRSM rows", file=sys.stderr, flush=True)
        other["strsm_4096x1024"] = strsm_rows(h, ib, stream, a.steps)
    if "ctrsm" in rows:
        print("[bench] CTRSM rows", file=sys.stderr, flush=True)
        other["ctrsm_4096x1024"] = ctrsm_rows(h, ib, stream, a.steps)
    if "chain" in rows:
        print("[bench] GEMM+TRSM chain row", file=sys.stderr, flush=True)
        other["chain_2048x512"] = chain_rows(h, ib, stream, a.steps)
    if "level3" in rows:
        print("[bench] level-3 rows", file=sys.stderr, flush=True)
        other["level3"] = level3_rows(h, ib, stream, a.steps)
    if "gemv" in rows:
        print("[bench] GEMV rows", file=sys.stderr, flush=True)
        other["gemv"] = gemv_rows(h, ib, stream, a.steps)
    if "host" in rows:
        print("[bench] host-pointer rows", file=sys.stderr, flush=True)
        other["host_pointer"] = host_rows(ib, a.steps)
    if other:
        line["other_rows"] = other
    if device:
        # The roofline peaks above are the published MI355X figures (2.4 GHz, 256 CUs);
        # these are the same peaks from this box's own agent properties.
        line["device"] = device
    if not a.no_cpu_baseline and A is not None:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(A, B, a.size, a.cpu_seconds)
    h.close()
    checks = row_checks(line)
    failed = [t for ok, t in checks if not ok]
    line["checks_failed"] = len(failed)
    print(json.dumps(line), flush=True)
    # compact per-configuration summary (stderr, after the JSON line)
    for _, text in checks:
        print("[bench] " + text, file=sys.stderr)
    cb = line.get("cpu_baseline")
    if cb:
        print(f"[bench] cpu_baseline: {cb['value'] * 1e3:.2f} GFLOP/s on {cb['cores']} threads, single thread "
              f"{cb['single_thread']['value'] * 1e3:.3f} GFLOP/s ({cb['single_thread']['kind']}), {cb['cpu_model']}",
              file=sys.stderr)
    sys.stderr.flush()
    if failed:
        print(f"[bench] {len(failed)} configuration(s) outside their error bound", file=sys.stderr, flush=True)
        sys.exit(1)


if __name__ == "__main__":
    main()
