"""Benchmark of the MI355X QR / polar path on one GPU.

Prints ONE JSON line (the bench contract).  The headline (`value`) is
qr_compact! on one 8192x1024 Float64 Gaussian matrix (positive = true): a step
is one qr_compact! call; every timed step factors its own copy of the input,
prepared in HBM before the timed region (qr_compact! destroys its input);
value = algorithmic TFLOP/s with LAPACK's flop counts (dgeqrf + dorgqr,
4 m n^2 - 4 n^3 / 3 per call).

The roofline object uses the same call as its launch unit (the library call is
a sequence of kernels; per-kernel times are in profiles/).  `traffic` is the
HBM bytes per call measured by rocprofv3 counters (FETCH_SIZE + WRITE_SIZE over
all kernels of one call, scaled by the copy-kernel calibration factor), passed
in with --traffic-bytes or read from profiles/qr_compact_8192x1024_f64_hbm.json
(FETCH_SIZE scaled by the copy calibration factor 2.0, WRITE_SIZE by 1.0).

`extras` holds the other measured workloads of the same path, each timed with
HIP events on the stream the library launches on:
  qr_full! 8192x1024 Float64, qr_compact!/qr_full! 8192x1024 ComplexF64,
  left_polar! 4096x4096 Float64 Gaussian (FP64 QDWH), with the host's
  dgesdd-route polar (the reference's PolarViaSVD) timed once beside it,
  left_polar! 4096x4096 with sigma log-spaced 1 .. 1e-6 in FP64 / FP32 / BF16
  bulk (speedups over FP64, W error against the FP64 GPU result and against
  the exact U V^T of the construction); polar TFLOP/s count 26n^3/3 per
  QR-form step, 10n^3/3 per Cholesky-form step, 3n^3 per Newton-Schulz or
  polynomial step and 2n^3 for P, from the reported iteration counts,
  qr and left-polar pullbacks on 256 x (2048 x 256) Float64,
  batched qr_compact! on 4096 blocks (m log-uniform in [16, 1024], n uniform in
  [8, m]) in Float64 and ComplexF64, one captured-graph execution per step.

cpu_baseline runs the reference's LAPACK sequence (dgeqrt nb=36 -> dgemqrt on
I -> triu + sign gauge, restated in oracle/mak_oracle.py through scipy's
LAPACK) on the host cores for a bounded sample of the headline workload, with
OPENBLAS_NUM_THREADS set to the physical cores of the CPU affinity mask
(capped by the host's OMP_NUM_THREADS share); `cores` is that thread count.
"""
from __future__ import annotations

import argparse
import json
import math
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
METRIC = "QR/polar TFLOP/s (% FP64 peak); mixed- vs FP64-polar speedup @1e-12 err"


def geqrf_flops(m, n):
    """LAPACK xGEQRF count (real flops)."""
    if m >= n:
        return 2.0 * m * n * n - 2.0 * n ** 3 / 3.0
    return 2.0 * n * m * m - 2.0 * m ** 3 / 3.0


def orgqr_flops(m, nq, k):
    """LAPACK xORGQR count: m x nq Q from k reflectors."""
    return 4.0 * m * nq * k - 2.0 * (m + nq) * k * k + 4.0 * k ** 3 / 3.0


def qr_flops(m, n, full=False, complex_=False):
    k = min(m, n)
    f = geqrf_flops(m, n) + orgqr_flops(m, m if full else k, k)
    return 4.0 * f if complex_ else f


def qr_pullback_flops(m, n):
    """Q re-formed from WY (orgqr) + Q^H X, Q M, X R^-H (3 x 2 m n^2) + R11 products / inverse."""
    return orgqr_flops(m, n, n) + 6.0 * m * n * n + 2.0 * n ** 3


def host_cores():
    """(physical cores in this process's CPU affinity mask, threads to use):
    logical CPUs are mapped to (package, core) through sysfs; the thread count
    is that, capped by OMP_NUM_THREADS when the host sets a share."""
    cpus = sorted(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else list(range(os.cpu_count()))
    cores = set()
    for c in cpus:
        topo = Path(f"/sys/devices/system/cpu/cpu{c}/topology")
        try:
            cores.add((topo.joinpath("physical_package_id").read_text().strip(),
                       topo.joinpath("core_id").read_text().strip()))
        except OSError:
            cores.add(("?", str(c)))
    phys = max(1, len(cores))
    share = os.environ.get("OMP_NUM_THREADS", "")
    threads = min(phys, int(share)) if share.isdigit() and int(share) > 0 else phys
    return phys, threads


def blas_threads():
    try:
        from threadpoolctl import threadpool_info
        return max((d.get("num_threads", 1) for d in threadpool_info() if d.get("user_api") == "blas"),
                   default=None)
    except Exception:  # pragma: no cover
        return None


def cpu_baseline(m, n, budget_s=12.0):
    import numpy as np
    from oracle import mak_oracle as O
    phys, _ = host_cores()
    threads = blas_threads() or int(os.environ.get("OPENBLAS_NUM_THREADS", "1"))
    A = O.gaussian(m, n, np.float64, seed=0)
    O.qr_compact(A)                       # warm-up
    times = []
    t_start = time.perf_counter()
    while len(times) < 3 or (time.perf_counter() - t_start < budget_s and len(times) < 50):
        t0 = time.perf_counter()
        O.qr_compact(A)
        times.append(time.perf_counter() - t0)
    t = float(np.median(times))
    return {"value": round(qr_flops(m, n) / t / 1e12, 4), "unit": "TFLOP/s", "cores": int(threads),
            "physical_cores_in_mask": phys, "kind": "port",
            "sample": f"{len(times)} x qr_compact 8192x1024 Float64 (median {t * 1e3:.1f} ms), "
                      "dgeqrt(nb=36)+dgemqrt on I via scipy LAPACK"}


def timed(fn, steps, warmup, stream):
    """Median ms per call of fn(i) with HIP events on `stream` (fn(i) for i < warmup untimed)."""
    import torch
    for i in range(warmup):
        fn(i)
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(steps)]
    for i in range(steps):
        ev[i][0].record(stream)
        fn(warmup + i)
        ev[i][1].record(stream)
    torch.cuda.synchronize()
    t = sorted(a.elapsed_time(b) for a, b in ev)
    return t[len(t) // 2] if len(t) % 2 else 0.5 * (t[len(t) // 2 - 1] + t[len(t) // 2])


def gaussian_dev(m, n, dtype, seed, dev):
    import torch
    g = torch.Generator(device="cpu").manual_seed(seed)
    if dtype == torch.complex128:
        x = torch.complex(torch.randn((n, m), generator=g, dtype=torch.float64),
                          torch.randn((n, m), generator=g, dtype=torch.float64))
    else:
        x = torch.randn((n, m), generator=g, dtype=torch.float64)
    return x.to(dev).t()          # column-major m x n


def bench_qr_extra(mak, dev, stream, m, n, dtype, full, steps):
    import torch
    A0 = gaussian_dev(m, n, dtype, 11, dev)
    copies = [A0.clone() for _ in range(steps + 1)]
    Q, R = mak.initialize_output_qr(A0, full)
    fn = (lambda i: mak.qr_full_(copies[i], (Q, R))) if full else (lambda i: mak.qr_compact_(copies[i], (Q, R)))
    ms = timed(fn, steps, 1, stream)
    res = (torch.linalg.norm(A0 - Q @ R) / torch.linalg.norm(A0)).item()
    k = Q.shape[1]
    orth = (torch.linalg.norm(Q.mH @ Q - torch.eye(k, dtype=dtype, device=dev)) / math.sqrt(k)).item()
    fl = qr_flops(m, n, full, dtype == torch.complex128)
    return {"ms": round(ms, 4), "tflops": round(fl / ms / 1e9, 3), "rel_residual": res, "orth_per_sqrt_n": orth}


def polar_flops(n, info):
    """Algorithmic flops of one square left_polar! call from its iteration
    counts: QR-form step 26n^3/3, Cholesky-form step 10n^3/3, Newton-Schulz
    and cubic polynomial steps 3n^3 (Gram n^3 + product 2n^3), P = W^H A 2n^3.
    Returns (total, {precision: flops}) with the share each precision runs in
    (reduced precision: the QR-form step's second-pass products (2n^3), the
    polynomial steps, the Newton-Schulz products (2n^3) and P run on FP32
    MFMA, BF16 for the late polynomial products at precision 16)."""
    n3 = float(n) ** 3
    q, c, ns, po = info["qr_steps"], info["chol_steps"], info["ns_steps"], info.get("poly_steps", 0)
    total = q * 26 * n3 / 3 + c * 10 * n3 / 3 + (ns + po) * 3 * n3 + 2 * n3
    if ns == 0 and po == 0:
        return total, {"f64": total}
    lo = q * 2 * n3 + po * 3 * n3 + ns * 2 * n3 + 2 * n3
    return total, {"f64": total - lo, "f32_or_bf16": lo}


def bench_polar(mak, dev, stream, steps, cpu=True):
    import numpy as np
    import torch
    n = 4096
    out = {}
    eye = torch.eye(n, dtype=torch.float64, device=dev)

    def props(A, W, P):
        return {"bwd": (torch.linalg.norm(A - W @ P) / torch.linalg.norm(A)).item(),
                "orth": (torch.linalg.norm(W.T @ W - eye) / math.sqrt(n)).item()}

    A = mak.colmajor(gaussian_dev(n, n, torch.float64, 21, dev))
    W, P = mak.initialize_output_polar(A, False)
    ms = timed(lambda i: mak.left_polar_(A, (W, P)), steps, 1, stream)
    info = mak.last_polar_info()
    fl, _ = polar_flops(n, info)
    rec = {"ms": round(ms, 3), "tflops": round(fl / ms / 1e9, 3),
           "frac_fp64_peak": round(fl / ms / 1e9 / FP64_PEAK_TFLOPS, 4), **props(A, W, P), "info": info}
    if cpu:
        # the reference's polar route on the host cores: dgesdd 'S' -> U V^H -> P
        from oracle import mak_oracle as O
        Ah = A.cpu().numpy()
        t0 = time.perf_counter()
        O.left_polar_svd(Ah)
        t_cpu = time.perf_counter() - t0
        rec["cpu_baseline"] = {"ms": round(t_cpu * 1e3, 1), "cores": blas_threads(), "kind": "port",
                               "sample": "1 x left_polar! 4096^2 Float64 via dgesdd (oracle left_polar_svd)",
                               "gpu_speedup": round(t_cpu * 1e3 / ms, 1)}
    out["left_polar_f64_gaussian_4096"] = rec

    # A = U diag(sigma) V^T, U, V Haar (positive-gauge QR of Gaussians), sigma 1 .. 1e-6
    U, _ = mak.qr_compact(gaussian_dev(n, n, torch.float64, 22, dev))
    V, _ = mak.qr_compact(gaussian_dev(n, n, torch.float64, 23, dev))
    sig = torch.logspace(0, -6, n, dtype=torch.float64, device=dev)
    A = mak.colmajor((U * sig[None, :]) @ V.T)
    Wx = U @ V.T
    res = {}
    for prec in (64, 32, 16):
        W, P = mak.initialize_output_polar(A, False)
        alg = mak.MI355XPolar(precision=prec)
        ms = timed(lambda i: mak.left_polar_(A, (W, P), alg), steps, 1, stream)
        res[prec] = (ms, W.clone(), props(A, W, P), mak.last_polar_info())
    W64 = res[64][1]
    nrm = torch.linalg.norm(W64).item()
    for prec, (ms, W, pr, info) in res.items():
        key = {64: "f64", 32: "f32_bulk", 16: "bf16_bulk"}[prec]
        fl, split = polar_flops(n, info)
        out[f"left_polar_{key}_cond1e6_4096"] = {
            "ms": round(ms, 3), "speedup_vs_f64": round(res[64][0] / ms, 3),
            "tflops": round(fl / ms / 1e9, 3), "gflop_by_precision": {k: round(v / 1e9, 1) for k, v in split.items()},
            "relW_vs_f64_gpu": (torch.linalg.norm(W - W64) / nrm).item(),
            "relW_vs_exact_UVt": (torch.linalg.norm(W - Wx) / nrm).item(), **pr,
            "iterations": info["iterations"], "qr_steps": info["qr_steps"], "chol_steps": info["chol_steps"],
            "poly_steps": info.get("poly_steps", 0), "ns_steps": info["ns_steps"]}
    return out


def bench_pullbacks(mak, dev, stream, steps, batch=256, m=2048, n=256):
    import torch
    out = {}
    tapes, Rs, Ws, Ps = [], [], [], []
    for b in range(batch):
        A = gaussian_dev(m, n, torch.float64, 1000 + b, dev)
        Q, R, tape = mak.qr_compact_rrule(A)
        tapes.append(tape)
        Rs.append(R)
        W, P = mak.left_polar(mak.colmajor(A))
        Ws.append(W)
        Ps.append(P)
    Awy = mak.colmajor_batch(torch.stack([t.Awy for t in tapes]))
    T = torch.stack([t.T for t in tapes])
    R3 = mak.colmajor_batch(torch.stack(Rs))
    del tapes, Rs
    W3 = mak.colmajor_batch(torch.stack(Ws))
    P3 = mak.colmajor_batch(torch.stack(Ps))
    del Ws, Ps
    g = torch.Generator(device="cpu").manual_seed(5)

    def cot(b, r, c):
        return mak.colmajor_batch(torch.randn((b, r, c), generator=g, dtype=torch.float64).to(dev))

    dQ, dR = cot(batch, m, n), torch.triu(cot(batch, n, n))
    dR = mak.colmajor_batch(dR)
    dA = mak.empty_colmajor_batch(batch, m, n, torch.float64, dev)
    ms = timed(lambda i: mak.qr_pullback_batched_(dA, Awy, T, R3, dQ, dR, p=n), steps, 1, stream)
    fl = batch * qr_pullback_flops(m, n)
    out["qr_pullback_256x2048x256"] = {"ms": round(ms, 3), "tflops": round(fl / ms / 1e9, 3)}
    del Awy, T, R3, dQ, dR
    dW, dP = cot(batch, m, n), cot(batch, n, n)
    iters = []

    def pol(i):
        _, it = mak.left_polar_pullback_batched_(dA, W3, P3, dW, dP)
        iters.append(it)
    ms = timed(pol, steps, 1, stream)
    out["left_polar_pullback_256x2048x256"] = {"ms": round(ms, 3), "cg_iterations": iters[-1]}
    return out


def bench_batched(mak, dev, stream, steps, dtype, count=4096):
    """4096 blocks, m log-uniform in [16, 1024], n uniform in [8, m], one
    captured-graph execution per step (inputs refilled before each, untimed)."""
    import torch
    shapes = mak.block_shapes(count, seed=0)
    g = torch.Generator(device="cpu").manual_seed(1)
    A0 = []
    for m, n in shapes:
        x = torch.randn((n, m), generator=g, dtype=torch.float64)
        if dtype == torch.complex128:
            x = torch.complex(x, torch.randn((n, m), generator=g, dtype=torch.float64))
        A0.append(x.to(dev).t())
    As = [a.clone() for a in A0]
    t0 = time.perf_counter()
    plan = mak.BatchedQRPlan(As, tune=True)
    t_plan = time.perf_counter() - t0
    cplx = dtype == torch.complex128
    fl = sum(qr_flops(m, n, False, cplx) for m, n in shapes)
    ev = []
    for i in range(steps + 1):
        for a, a0 in zip(As, A0):
            a.copy_(a0)
        e = (torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True))
        e[0].record(stream)
        plan.execute()
        e[1].record(stream)
        ev.append(e)
    torch.cuda.synchronize()
    t = sorted(a.elapsed_time(b) for a, b in ev[1:])
    ms = t[len(t) // 2]
    Q, R = plan.Qs[0], plan.Rs[0]
    res = (torch.linalg.norm(A0[0] - Q @ R) / torch.linalg.norm(A0[0])).item()
    info = plan.info()
    plan.close()
    return {"ms": round(ms, 3), "tflops": round(fl / ms / 1e9, 3), "gflop": round(fl / 1e9, 2),
            "blocks": count, "plan_s": round(t_plan, 3), "plan": info, "rel_residual_block0": res}


def run_extras(mak, dev, stream, groups, xs, cpu=True):
    import torch
    m, n = 8192, 1024
    ex = {}
    if "qr" in groups:
        ex["qr_full_f64_8192x1024"] = bench_qr_extra(mak, dev, stream, m, n, torch.float64, True, xs)
        ex["qr_compact_c64_8192x1024"] = bench_qr_extra(mak, dev, stream, m, n, torch.complex128, False, xs)
        ex["qr_full_c64_8192x1024"] = bench_qr_extra(mak, dev, stream, m, n, torch.complex128, True, xs)
        torch.cuda.empty_cache()
    if "polar" in groups:
        ex.update(bench_polar(mak, dev, stream, xs, cpu))
        torch.cuda.empty_cache()
    if "pullback" in groups:
        ex.update(bench_pullbacks(mak, dev, stream, xs))
        torch.cuda.empty_cache()
    if "batched" in groups:
        ex["qr_batched_f64_4096_blocks"] = bench_batched(mak, dev, stream, xs, torch.float64)
        torch.cuda.empty_cache()
        ex["qr_batched_c64_4096_blocks"] = bench_batched(mak, dev, stream, xs, torch.complex128)
    return ex


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline leg")
    ap.add_argument("--no-extras", action="store_true", help="headline workload only")
    ap.add_argument("--extras", default="qr,polar,pullback,batched",
                    help="comma-separated extra workload groups: qr, polar, pullback, batched")
    ap.add_argument("--skip-headline", action="store_true",
                    help="profiling aid: run only the extra workloads (the JSON line then has no headline value)")
    ap.add_argument("--traffic-bytes", type=float, default=None,
                    help="HBM bytes per qr_compact! call from the rocprofv3 counter runs (default: the "
                         "committed counter summary in profiles/, if present)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run a 1 GiB device copy (the FETCH_SIZE / WRITE_SIZE calibration kernel)")
    ap.add_argument("--polar-profile", type=int, default=None, metavar="PREC",
                    help="profiling aid: only left_polar! on the 4096^2 cond-1e6 input at precision PREC "
                         "(64 / 32 / 16), --warmup untimed + --steps timed calls")
    args = ap.parse_args()
    if args.traffic_bytes is None:
        prof = ROOT / "profiles" / "qr_compact_8192x1024_f64_hbm.json"
        if prof.exists():
            args.traffic_bytes = json.loads(prof.read_text()).get("traffic_bytes_per_call")
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (got --gpus {args.gpus}); the QR/polar path "
              "runs on a single MI355X in this build", file=sys.stderr)
        sys.exit(2)

    # host BLAS threads for the CPU legs: the physical cores of the affinity
    # mask (capped by the host's OMP_NUM_THREADS share), set before numpy loads
    _, threads = host_cores()
    os.environ.setdefault("OPENBLAS_NUM_THREADS", str(threads))
    import numpy as np
    import torch
    import mak_mi355x as mak

    dev = torch.device("cuda:0")
    stream = torch.cuda.current_stream(dev)
    if args.copy_calibration:
        src = torch.ones(1 << 27, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        for _ in range(args.steps):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_each_way": src.numel() * 8, "launches": args.steps}), flush=True)
        return

    if args.polar_profile is not None:
        n = 4096
        U, _ = mak.qr_compact(gaussian_dev(n, n, torch.float64, 22, dev))
        V, _ = mak.qr_compact(gaussian_dev(n, n, torch.float64, 23, dev))
        sig = torch.logspace(0, -6, n, dtype=torch.float64, device=dev)
        A = mak.colmajor((U * sig[None, :]) @ V.T)
        W, P = mak.initialize_output_polar(A, False)
        alg = mak.MI355XPolar(precision=args.polar_profile)
        ms = timed(lambda i: mak.left_polar_(A, (W, P), alg), args.steps, args.warmup, stream)
        info = mak.last_polar_info()
        fl, split = polar_flops(n, info)
        print(json.dumps({"workload": f"left_polar! 4096^2 cond 1e6 precision {args.polar_profile}",
                          "ms": round(ms, 3), "tflops": round(fl / ms / 1e9, 3),
                          "gflop_by_precision": {k: round(v / 1e9, 1) for k, v in split.items()},
                          "info": info}), flush=True)
        return

    m, n = 8192, 1024
    groups = set() if args.no_extras else {g for g in args.extras.split(",") if g}
    if args.skip_headline:
        ex = run_extras(mak, dev, stream, groups, max(2, args.steps // 3), not args.no_cpu)
        print(json.dumps({"metric": METRIC, "value": None, "extras": ex}), flush=True)
        return
    rng = np.random.default_rng(0)
    A0 = torch.from_numpy(np.asfortranarray(rng.standard_normal((m, n)))).to(dev)
    A0 = mak.colmajor(A0)
    nsteps = args.steps + args.warmup
    copies = [A0.clone() for _ in range(nsteps)]
    Q, R = mak.initialize_output_qr(A0, False)
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
    ms = float(np.median(per_call_ms))
    tflops = qr_flops(m, n) / (ms * 1e-3) / 1e12
    del copies

    # correctness guard on the last step (cheap properties, on the device)
    res = (torch.linalg.norm(A0 - Q @ R) / torch.linalg.norm(A0)).item()
    orth = torch.linalg.norm(Q.T @ Q - torch.eye(n, dtype=Q.dtype, device=dev)).item()

    out = {
        "metric": METRIC,
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
                   "parallelism": "single-gpu",
                   "precision": "FP64 factors (CholeskyQR2 on FP64 MFMA); the final correction term Q1 D "
                                "(||D|| ~ 1e-14) on BF16 MFMA with FP32 accumulation, its error bound below "
                                "FP64 rounding; checks holds the FP64 residual and orthogonality"},
        "roofline": {"bound": "mfma", "achieved": round(tflops, 3), "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(tflops / FP64_PEAK_TFLOPS, 4),
                     "traffic": args.traffic_bytes,
                     "launch_unit": "one qr_compact! call (median of event-timed calls on its stream)",
                     "ms_per_launch": round(ms, 4)},
        "checks": {"rel_residual": res, "orth": orth},
    }
    if groups:
        out["extras"] = run_extras(mak, dev, stream, groups, max(2, args.steps // 3), not args.no_cpu)
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(m, n)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
