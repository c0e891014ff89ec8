"""Benchmark: qr_compact! on 8192x1024 Float64 (positive = true) on one MI355X.

Prints ONE JSON line (the bench contract).  A step is one qr_compact! call on
one 8192x1024 Float64 Gaussian matrix; every timed step factors its own copy
of the input, prepared in HBM before the timed region (qr_compact! destroys
its input).  value = algorithmic TFLOP/s (4 m n^2 - 4 n^3 / 3 per call).

The roofline object uses the same call as its launch unit (the library call is
a sequence of kernels; per-kernel times are in profiles/).  cpu_baseline runs
the reference's LAPACK sequence (dgeqrt nb=36 -> dgemqrt on I -> triu + sign
gauge, restated in oracle/mak_oracle.py through scipy's LAPACK) on the host
cores for a bounded sample of the same workload.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
for p in (ROOT, ROOT / "matrixalgebrakit.jl_amd"):
    if str(p) not in sys.path:
        sys.path.insert(0, str(p))

FP64_PEAK_TFLOPS = 78.6   # MI355X dense FP64 MFMA
HBM_PEAK_GBS = 8000.0


def qr_flops(m, n):
    return 4.0 * m * n * n - 4.0 * n ** 3 / 3.0


def cpu_baseline(m, n, budget_s=12.0):
    import numpy as np
    from oracle import mak_oracle as O
    try:
        from threadpoolctl import threadpool_info
        threads = max((d.get("num_threads", 1) for d in threadpool_info()
                       if d.get("user_api") == "blas"), default=os.cpu_count())
    except Exception:  # pragma: no cover
        threads = os.cpu_count()
    A = O.gaussian(m, n, np.float64, seed=0)
    O.qr_compact(A)                       # warm-up
    times = []
    t_start = time.perf_counter()
    while len(times) < 3 or (time.perf_counter() - t_start < budget_s and len(times) < 50):
        t0 = time.perf_counter()
        O.qr_compact(A)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": qr_flops(m, n) / t / 1e12, "unit": "TFLOP/s", "cores": int(threads),
            "kind": "port",
            "sample": f"{len(times)} x qr_compact 8192x1024 Float64 (median {t * 1e3:.1f} ms), "
                      "dgeqrt(nb=36)+dgemqrt on I via scipy LAPACK"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the QR/polar path "
              "runs on a single MI355X in this build", file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    import mak_mi355x as mak

    dev = torch.device("cuda:0")
    m, n = 8192, 1024
    rng = np.random.default_rng(0)
    A0 = torch.from_numpy(np.asfortranarray(rng.standard_normal((m, n)))).to(dev)
    A0 = mak.colmajor(A0)
    nsteps = args.steps + args.warmup
    copies = [A0.clone() for _ in range(nsteps)]
    Q, R = mak.initialize_output_qr(A0, False)
    stream = torch.cuda.current_stream(dev)
    alg = mak.MI355XHouseholder(positive=True)
    for i in range(args.warmup):
        mak.qr_compact_(copies[i], (Q, R), alg)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
          for _ in range(args.steps)]
    t0 = time.perf_counter()
    for i in range(args.steps):
        ev[i][0].record(stream)
        mak.qr_compact_(copies[args.warmup + i], (Q, R), alg)
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    per_call_ms = [a.elapsed_time(b) for a, b in ev]
    ms = float(np.mean(per_call_ms))
    tflops = qr_flops(m, n) / (ms * 1e-3) / 1e12

    # correctness guard on the last step (cheap properties, on the device)
    Af = A0
    res = (torch.linalg.norm(Af - Q @ R) / torch.linalg.norm(Af)).item()
    orth = torch.linalg.norm(Q.T @ Q - torch.eye(n, dtype=Q.dtype, device=dev)).item()

    out = {
        "metric": "QR/polar TFLOP/s (% FP64 peak); mixed- vs FP64-polar speedup @1e-12 err",
        "value": round(tflops, 3),
        "unit": "TFLOP/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "qr_compact! 8192x1024 Float64 Gaussian, positive=true", "m": m, "n": n,
                   "parallelism": "single-gpu"},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 3), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(tflops / FP64_PEAK_TFLOPS, 4), "traffic": None,
                     "launch_unit": "one qr_compact! call (event-timed on its stream)",
                     "ms_per_launch": round(ms, 4)},
        "checks": {"rel_residual": res, "orth": orth},
    }
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(m, n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
