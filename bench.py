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
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--size", type=int, default=8192)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0, help="target CPU-baseline sample duration")
    ap.add_argument("--paths", default="fp32,bf16x3")
    ap.add_argument("--impl", default="", help="comma list, one implementation name per path (empty = automatic)")
    ap.add_argument("--trans", default="NN", help="NN, NT, TN or TT for the timed op")
    ap.add_argument("--no-check", action="store_true", help="skip the FP64 error check (profiling runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="run one 1 GiB device-to-device copy kernel first (scales HBM byte counters)")
    return ap.parse_args()


def cpu_baseline(A: np.ndarray, B: np.ndarray, n: int, target_s: float) -> dict:
    """Times the restated reference CPU GEMM (oracle.sgemm_mt = cpu_gemm split over
    output columns) on a slab of columns of the same workload."""
    import oracle
    m = k = A.shape[0]
    threads = max(1, min(os.cpu_count() or 1, 16))
    cols = threads
    while True:
        C = np.zeros((m, cols), np.float32, order="F")
        Bs = np.asfortranarray(B[:, :cols])
        t0 = time.perf_counter()
        oracle.sgemm_mt(threads, 0, 0, m, cols, k, 1.0, A, Bs, 0.0, C)
        dt = time.perf_counter() - t0
        if dt >= target_s or cols >= n:
            break
        cols = min(n, max(cols + threads, int(cols * min(8.0, 1.2 * target_s / max(dt, 1e-3)))))
        cols = max(threads, (cols // threads) * threads)
    flops = 2.0 * m * cols * k
    return {"value": flops / dt / 1e12, "unit": "TFLOP/s", "cores": threads, "kind": "port",
            "sample": f"SGEMM NN {m}x{cols}x{k} column slab of the {m}x{n}x{k} workload "
                      f"({cols} of {n} output columns, {dt:.1f} s), naive cpu_gemm loop order with a double "
                      f"accumulator, columns split over {threads} threads; rate extrapolates linearly in n"}


def main():
    a = parse()
    if a.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {a.gpus} is not supported (this build has no multi-GPU path)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from clgpu_amd import datagen
    from clgpu_amd import iclblas as ib

    n = a.size
    m = k = n
    ta, tb = ({"N": 0, "T": 1}[a.trans[0]], {"N": 0, "T": 1}[a.trans[1]])
    print(f"[bench] generating {n}x{n} inputs", file=sys.stderr, flush=True)
    A = datagen.matrix(n, n, name="A", seed=0)
    B = datagen.matrix(n, n, name="B", seed=0)
    dev = torch.device("cuda:0")
    torch.cuda.set_device(dev)
    # column-major (ld x cols) <-> torch (cols, ld) row-major
    tA = torch.from_numpy(np.ascontiguousarray(A.T)).to(dev)
    tB = torch.from_numpy(np.ascontiguousarray(B.T)).to(dev)
    tC = torch.empty((n, n), device=dev, dtype=torch.float32)
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
        start = torch.cuda.Event(enable_timing=True)
        stop = torch.cuda.Event(enable_timing=True)
        start.record(stream)
        for _ in range(a.steps):
            step()
        stop.record(stream)
        torch.cuda.synchronize()
        ms = start.elapsed_time(stop) / a.steps
        impl = h.last_implementation()
        tflops = flops / (ms * 1e-3) / 1e12
        peak = PEAK_FP32_MFMA if path == "fp32" else PEAK_BF16_MFMA
        executed = tflops if path == "fp32" else 3.0 * tflops
        results[path] = {"implementation": impl, "ms": ms, "tflops": tflops,
                         "frac_of_fp32_mfma_peak": tflops / PEAK_FP32_MFMA,
                         "executed_tflops": executed, "mfma_peak": peak, "mfma_frac": executed / peak}
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
    traffic = None
    tf = ROOT / "profiles" / "hbm_traffic.json"
    if tf.exists():
        traffic = json.loads(tf.read_text()).get(hr["implementation"])
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
                     "kernel": hr["implementation"],
                     "note": "achieved = executed MFMA flops per launch (3 x 2mnk for split-BF16) / mean launch time"},
    }
    if "fp32" in results:
        r = results["fp32"]
        t32 = None
        if tf.exists():
            t32 = json.loads(tf.read_text()).get(r["implementation"])
        line["roofline_fp32"] = {"bound": "mfma", "achieved": r["tflops"], "peak": PEAK_FP32_MFMA,
                                 "unit": "TFLOP/s", "frac": r["tflops"] / PEAK_FP32_MFMA, "traffic": t32,
                                 "kernel": r["implementation"]}
        if "bf16x3" in results:
            line["split_over_fp32"] = results["bf16x3"]["tflops"] / r["tflops"]
    if not a.no_cpu_baseline:
        print("[bench] CPU baseline", file=sys.stderr, flush=True)
        line["cpu_baseline"] = cpu_baseline(A, B, n, a.cpu_seconds)
    h.close()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
