#!/usr/bin/env python3
"""Headline benchmark: FP64 TSMTTSM and TSMM throughput on one MI355X.

One step = the metric's sweep: for M = N in {1, 2, 4, 8, 16, 32, 64},
TSMTTSM C = A^T B (A, B row-major K x M, K = 2^26) followed by TSMM
C = A B (A row-major K x M, B M x M), alpha = 1, beta = 0, out of place,
14 calls on inputs already resident in HBM (generated on the device from the
counter RNG, U[-1,1), seed 1).  Bytes are the reference's algorithmic counts
(perf.cu:209-231):  TSMTTSM (M + N) K 8,  TSMM (M K + N K + M N) 8.

Prints ONE JSON line:  value = total algorithmic bytes of the timed steps /
wall time (GB/s); per-width GB/s from HIP events on the launch stream;
roofline of the dominant (longest) kernel against the MI355X HBM peak; and a
CPU baseline (the oracle's OpenMP FP64 loops, timed on this host on a bounded
sample of the same sweep).

    python bench.py [--gpus 1] [--steps 10] [--warmup 2] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

WIDTHS = (1, 2, 4, 8, 16, 32, 64)
HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    sys.stderr.write(msg + "\n")
    sys.stderr.flush()


def op_bytes(op: str, M: int, N: int, K: int) -> int:
    if op == "tsmttsm":
        return (M + N) * K * 8
    return (M * K + N * K + M * N) * 8


def traffic_from_profiles(kernel_key: str):
    """HBM bytes per launch measured with rocprofv3 PMC counters, if a
    committed summary for this kernel exists (profiles/traffic.json)."""
    p = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(p) as f:
            d = json.load(f)
        v = d.get(kernel_key)
        return None if v is None else float(v["hbm_bytes_per_launch"])
    except (OSError, ValueError, KeyError):
        return None


def cpu_baseline(K_cpu: int, min_seconds: float, threads: int):
    import numpy as np

    import oracle

    oracle.build()
    Mmax = max(WIDTHS)
    A = oracle.uniform(K_cpu * Mmax, 1, 0)
    B = oracle.uniform(K_cpu * Mmax, 1, 1)
    C = np.zeros(Mmax * Mmax)
    T = np.zeros(K_cpu * Mmax)
    Bs = {M: oracle.uniform(M * M, 1, 2 + M) for M in WIDTHS}
    total_bytes = 0
    t0 = time.perf_counter()
    sweeps = 0
    while True:
        for M in WIDTHS:
            oracle.tsmttsm_par(threads, M, M, K_cpu, A, M, 1.0, B, M, 0.0, C, M)
            total_bytes += op_bytes("tsmttsm", M, M, K_cpu)
            oracle.tsmm_par(threads, M, M, K_cpu, A, M, 1.0, Bs[M], M, 0.0, T, M)
            total_bytes += op_bytes("tsmm", M, M, K_cpu)
        sweeps += 1
        el = time.perf_counter() - t0
        if el >= min_seconds:
            break
    return {
        "value": round(total_bytes / el / 1e9, 3),
        "unit": "GB/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"{sweeps} sweep(s) of the same 14 FP64 calls at K=2^{K_cpu.bit_length() - 1} "
                   f"(OpenMP oracle loops, {threads} threads), {el:.1f} s"),
    }


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.split("\n\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--logK", type=int, default=26)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-tune", action="store_true",
                    help="skip the untimed autotuning of each shape")
    args = ap.parse_args()

    if args.gpus != 1:
        log("bench.py runs on one GPU only: the tall-skinny products here are "
            "single-device kernels (K-sharding across GPUs is described but not "
            "built); rerun with --gpus 1")
        return 2

    import torch

    import skinny_matmul_amd as sk

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    sk.lib()
    dev = torch.device("cuda:0")
    K = 1 << args.logK
    Mmax = max(WIDTHS)
    log(f"allocating 2 x {K * Mmax * 8 / 2**30:.0f} GiB")
    A = torch.empty(K * Mmax, dtype=torch.float64, device=dev)
    B = torch.empty(K * Mmax, dtype=torch.float64, device=dev)
    Cs = torch.zeros(Mmax * Mmax, dtype=torch.float64, device=dev)
    sk.fill_uniform(A, A.numel(), 1, 0)
    sk.fill_uniform(B, B.numel(), 1, 1)
    Bsm = {}
    for M in WIDTHS:
        Bsm[M] = torch.empty(M * M, dtype=torch.float64, device=dev)
        sk.fill_uniform(Bsm[M], M * M, 1, 2 + M)
    torch.cuda.synchronize()

    ops = []
    for M in WIDTHS:
        ops.append(("tsmttsm", M))
        ops.append(("tsmm", M))

    def run(op, M):
        if op == "tsmttsm":
            ok = sk.dtsmttsm(0, M, M, K, A, M, 1.0, B, M, 0.0, Cs, M)
        else:
            ok = sk.dtsmm(0, M, M, K, A, M, 1.0, Bsm[M], M, 0.0, B, M)
        assert ok, (op, M)

    step_bytes = sum(op_bytes(op, M, M, K) for op, M in ops)
    tuned = {}
    if not args.no_tune:
        # untimed: the library's autotuner picks unroll and grid per shape
        for op, M in ops:
            tuned[f"{op}_{M}"] = round(sk.autotune(op, M, M, K, reps=3), 3)
        log(f"autotuned: {sk.tune_export().strip()}")
    for _ in range(args.warmup):
        for op, M in ops:
            run(op, M)
    torch.cuda.synchronize()
    log("warmup done")

    evs = {o: [] for o in ops}
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for o in ops:
            e0 = torch.cuda.Event(enable_timing=True)
            e1 = torch.cuda.Event(enable_timing=True)
            e0.record()
            run(*o)
            e1.record()
            evs[o].append((e0, e1))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0

    per = {}
    for o in ops:
        ms = sum(a.elapsed_time(b) for a, b in evs[o]) / len(evs[o])
        per[o] = ms
    per_width = {
        op: {str(M): round(op_bytes(op, M, M, K) / (per[(op, M)] * 1e-3) / 1e9, 1)
             for (o2, M) in ops if o2 == op}
        for op in ("tsmttsm", "tsmm")
    }
    dom = max(ops, key=lambda o: per[o])
    dom_bytes = op_bytes(dom[0], dom[1], dom[1], K)
    achieved = dom_bytes / (per[dom] * 1e-3) / 1e9
    kernel_key = f"{dom[0]}_f64_M{dom[1]}_K2^{args.logK}"
    traffic = traffic_from_profiles(kernel_key)
    value = step_bytes * args.steps / wall / 1e9
    log(f"per-width GB/s: {per_width}")

    cpu = None
    if not args.no_cpu:
        log("cpu baseline ...")
        cpu = cpu_baseline(1 << 22, args.cpu_seconds, min(16, os.cpu_count() or 1))

    out = {
        "metric": "TSMTTSM/TSMM GB/s (% HBM roofline) fp64 M=N 1-64",
        "value": round(value, 1),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic U[-1,1) from a counter RNG (seed 1), generated in HBM",
        "config": {
            "workload": (f"TSMTTSM C=A^T B then TSMM C=A B, fp64, M=N in {list(WIDTHS)}, "
                         f"K=2^{args.logK}, alpha=1, beta=0, out of place, "
                         f"{'default' if args.no_tune else 'autotuned'} blockCount"),
            "K": K,
            "widths": list(WIDTHS),
            "calls_per_step": len(ops),
            "bytes_per_step": step_bytes,
        },
        "per_width_GBs": per_width,
        "roofline": {
            "kernel": kernel_key,
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
        },
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))
    sys.stdout.flush()
    return 0


if __name__ == "__main__":
    sys.exit(main())
